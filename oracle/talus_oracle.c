/*
 * talus_oracle.c -- FP64 CPU restatement of the Talus deformable tet step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * library never links or calls it.
 *
 * What each part restates (reference paths relative to the Talus tree):
 *   mesh build      src/talus/deformable/solver/solver.cc:17-24 (KUHN),
 *                   76-78 (cornerOffset), 163-180 (supported), 2185-2418
 *   Lame moduli     src/talus/deformable/factor/factor.cc:21-30 (preciseLame)
 *   polar rotation  src/talus/deformable/material/material.cc:168-197
 *   inverted frame  material.cc:24-64 (elementRotation, Invertible mode)
 *   corotated tet   material.cc:136-164, 236-265 (tetStrain, trial)
 *   J2 return map   material.cc:594-630, 978-1031 (returnMap, tetPointTrial)
 *   SNH law         src/talus/dynamics/neohookean/neohookean.cc:41-47,
 *                   183-196; material.cc:886-946 (H-form first Piola)
 *   Hencky flow     src/talus/dynamics/strainstep/plasticflow.cc:96-260
 *   tear stretch    src/talus/dynamics/strainstep/tearing.cc:248-290, 469-483
 *   operator        solver.cc:904-1048 (applySystem)
 *   block Jacobi    material.cc:290-304, solver.cc:1574-1593
 *   PCG             solver.cc:1904-2089 (conjugateGradient)
 *   step            solver.cc:531-623, 3472-3950, 4351-4368, 5829-5849,
 *                   1260-1346 (commitPlastic)
 *
 * Where the restatement deliberately differs from the reference, the
 * difference is documented at the function.  Everything is FP64 and
 * deterministic: element results go to per-(tet,corner) slots and nodes sum
 * their slots in fixed incidence order, and dot products are summed in fixed
 * blocks, so results do not depend on the thread count.
 */
#include "talus_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* 3x3 helpers, row-major                                               */
/* ------------------------------------------------------------------ */

double orc_det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

/* cofactor matrix: cof(A) = det(A) A^-T */
static void cof3(const double *a, double *c) {
  c[0] = a[4] * a[8] - a[5] * a[7];
  c[1] = a[5] * a[6] - a[3] * a[8];
  c[2] = a[3] * a[7] - a[4] * a[6];
  c[3] = a[2] * a[7] - a[1] * a[8];
  c[4] = a[0] * a[8] - a[2] * a[6];
  c[5] = a[1] * a[6] - a[0] * a[7];
  c[6] = a[1] * a[5] - a[2] * a[4];
  c[7] = a[2] * a[3] - a[0] * a[5];
  c[8] = a[0] * a[4] - a[1] * a[3];
}

static int inv3(const double *a, double *out) {
  double c[9];
  cof3(a, c);
  const double d = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
  if (d == 0.0) return 0;
  const double s = 1.0 / d;
  /* A^-1 = cof(A)^T / det */
  out[0] = c[0] * s; out[1] = c[3] * s; out[2] = c[6] * s;
  out[3] = c[1] * s; out[4] = c[4] * s; out[5] = c[7] * s;
  out[6] = c[2] * s; out[7] = c[5] * s; out[8] = c[8] * s;
  return 1;
}

static void mul3(const double *a, const double *b, double *c) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      t[r * 3 + k] = a[r * 3 + 0] * b[0 * 3 + k] + a[r * 3 + 1] * b[1 * 3 + k] +
                     a[r * 3 + 2] * b[2 * 3 + k];
  memcpy(c, t, sizeof t);
}

/* c = a^T b */
static void mul3_tn(const double *a, const double *b, double *c) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      t[r * 3 + k] = a[0 * 3 + r] * b[0 * 3 + k] + a[1 * 3 + r] * b[1 * 3 + k] +
                     a[2 * 3 + r] * b[2 * 3 + k];
  memcpy(c, t, sizeof t);
}

/* c = a b^T */
static void mul3_nt(const double *a, const double *b, double *c) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      t[r * 3 + k] = a[r * 3 + 0] * b[k * 3 + 0] + a[r * 3 + 1] * b[k * 3 + 1] +
                     a[r * 3 + 2] * b[k * 3 + 2];
  memcpy(c, t, sizeof t);
}

static void mv3(const double *a, const double *v, double *o) {
  const double x = v[0], y = v[1], z = v[2];
  o[0] = a[0] * x + a[1] * y + a[2] * z;
  o[1] = a[3] * x + a[4] * y + a[5] * z;
  o[2] = a[6] * x + a[7] * y + a[8] * z;
}

static void mtv3(const double *a, const double *v, double *o) {
  const double x = v[0], y = v[1], z = v[2];
  o[0] = a[0] * x + a[3] * y + a[6] * z;
  o[1] = a[1] * x + a[4] * y + a[7] * z;
  o[2] = a[2] * x + a[5] * y + a[8] * z;
}

static double trace3(const double *a) { return a[0] + a[4] + a[8]; }

static double contract3(const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < 9; ++i) s += a[i] * b[i];
  return s;
}

static void ident3(double *a) {
  memset(a, 0, 9 * sizeof(double));
  a[0] = a[4] = a[8] = 1.0;
}

/* shape gradients g0..g3 from the rows of Dm^-1 (material.cc:199-207) */
static void shape_gradients(const double *dminv, double g[4][3]) {
  for (int c = 0; c < 3; ++c) {
    g[1][c] = dminv[0 * 3 + c];
    g[2][c] = dminv[1 * 3 + c];
    g[3][c] = dminv[2 * 3 + c];
    g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
  }
}

/* displacement gradient H = [u1-u0, u2-u0, u3-u0] Dm^-1 */
static void displacement_gradient(const double *u, const double *dminv,
                                  double *H) {
  double d[3][3]; /* d[k] = u_{k+1} - u_0 */
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) d[k][r] = u[(k + 1) * 3 + r] - u[r];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      H[r * 3 + c] = d[0][r] * dminv[0 * 3 + c] + d[1][r] * dminv[1 * 3 + c] +
                     d[2][r] * dminv[2 * 3 + c];
}

/* ------------------------------------------------------------------ */
/* Symmetric 3x3 eigen decomposition: cyclic Jacobi to full precision.  */
/* values ascending; vectors row-major with eigenvector i in column i.  */
/* Stands in for Eigen::SelfAdjointEigenSolver (plasticflow.cc:97,111;  */
/* tearing.cc:278), whose results are unique up to eigenvector sign.    */
/* ------------------------------------------------------------------ */
void orc_eigen_sym3(const double *A, double *values, double *vectors) {
  double a[9];
  memcpy(a, A, sizeof a);
  a[1] = a[3] = 0.5 * (A[1] + A[3]);
  a[2] = a[6] = 0.5 * (A[2] + A[6]);
  a[5] = a[7] = 0.5 * (A[5] + A[7]);
  double v[9];
  ident3(v);
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
    const double diag = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
    if (off <= 1e-36 * diag || off == 0.0) break;
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int k = 0; k < 3; ++k) {
      const int p = P[k], q = Q[k];
      const double apq = a[p * 3 + q];
      if (apq == 0.0) continue;
      const double app = a[p * 3 + p], aqq = a[q * 3 + q];
      const double theta = (aqq - app) / (2.0 * apq);
      const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                       (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
      /* A' = J^T A J with J the (p,q) rotation */
      for (int r = 0; r < 3; ++r) {
        const double arp = a[r * 3 + p], arq = a[r * 3 + q];
        a[r * 3 + p] = c * arp - s * arq;
        a[r * 3 + q] = s * arp + c * arq;
      }
      for (int r = 0; r < 3; ++r) {
        const double apr = a[p * 3 + r], aqr = a[q * 3 + r];
        a[p * 3 + r] = c * apr - s * aqr;
        a[q * 3 + r] = s * apr + c * aqr;
      }
      for (int r = 0; r < 3; ++r) {
        const double vrp = v[r * 3 + p], vrq = v[r * 3 + q];
        v[r * 3 + p] = c * vrp - s * vrq;
        v[r * 3 + q] = s * vrp + c * vrq;
      }
    }
  }
  double w[3] = {a[0], a[4], a[8]};
  int order[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (w[order[j]] < w[order[i]]) {
        const int t = order[i];
        order[i] = order[j];
        order[j] = t;
      }
  for (int i = 0; i < 3; ++i) {
    values[i] = w[order[i]];
    for (int r = 0; r < 3; ++r) vectors[r * 3 + i] = v[r * 3 + order[i]];
  }
}

/* ------------------------------------------------------------------ */
/* Materials                                                            */
/* ------------------------------------------------------------------ */

/* preciseLame (factor.cc:21-30): nu clamped to [0, 0.49], FP64. */
void orc_lame(float youngs, float poisson, double *lambda, double *mu) {
  double nu = (double)poisson;
  if (!(nu > 0.0)) nu = 0.0;
  if (nu > 0.49) nu = 0.49;
  const double e = (double)youngs;
  *lambda = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  *mu = e / (2.0 * (1.0 + nu));
}

static const orc_material *lane_of(const orc_material *materials,
                                   uint32_t lanes, uint32_t lane,
                                   orc_material *fallback) {
  /* laneOf (solver.cc:30-37): out-of-range lane -> default material */
  if (lane < lanes) return &materials[lane];
  memset(fallback, 0, sizeof *fallback);
  fallback->friction = 0.5f;
  fallback->density = 1000.0f;
  fallback->strength = INFINITY;
  fallback->youngs_modulus = 0.0f;
  fallback->poissons_ratio = 0.3f;
  fallback->yield_stress = INFINITY;
  fallback->elastic_model = 0u;
  fallback->damping_ratio = INFINITY;
  fallback->tear_strain = INFINITY;
  fallback->plastic_max = INFINITY;
  fallback->crush_stress = INFINITY;
  fallback->crush_limit = 1.0f;
  return fallback;
}

/* yields (material.cc:752-754): finite and positive yield stress */
static int yields(float yield) { return isfinite(yield) && yield > 0.0f; }

/* ------------------------------------------------------------------ */
/* Mesh build (buildMesh for fully solid lattice cells)                 */
/* ------------------------------------------------------------------ */

static const uint32_t KUHN[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7},
                                    {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};

int orc_build_lattice(const orc_lattice *L, const orc_material *materials,
                      uint32_t lanes, const orc_support *supports,
                      uint32_t nsupports, orc_mesh *out) {
  memset(out, 0, sizeof *out);
  if (L->res[0] < 2 || L->res[1] < 2 || L->res[2] < 2) return -1;
  const uint32_t cx = L->res[0] - 1, cy = L->res[1] - 1, cz = L->res[2] - 1;
  const size_t samples = (size_t)L->res[0] * L->res[1] * L->res[2];
  uint32_t *node_of = (uint32_t *)malloc(samples * sizeof(uint32_t));
  for (size_t s = 0; s < samples; ++s) node_of[s] = UINT32_MAX;
  uint32_t nodes = 0;
  size_t solid_cells = 0;
  /* first-touch node numbering in z, y, x cell scan, corners 0..7 */
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        const size_t cell = x + (size_t)cx * (y + (size_t)cy * z);
        if (L->solid && !L->solid[cell]) continue;
        ++solid_cells;
        for (uint32_t c = 0; c < 8; ++c) {
          const uint32_t ax = x + (c & 1u), ay = y + ((c >> 1) & 1u),
                         az = z + ((c >> 2) & 1u);
          const size_t s = ax + (size_t)L->res[0] * (ay + (size_t)L->res[1] * az);
          if (node_of[s] == UINT32_MAX) node_of[s] = nodes++;
        }
      }
  if (nodes == 0) {
    free(node_of);
    return -1;
  }
  out->nodes = nodes;
  out->rest = (double *)calloc(3 * (size_t)nodes, sizeof(double));
  out->mass = (double *)calloc(nodes, sizeof(double));
  out->anchored = (uint8_t *)calloc(nodes, 1);
  out->tet = (uint32_t *)malloc(4 * 6 * solid_cells * sizeof(uint32_t));
  out->dminv = (double *)malloc(9 * 6 * solid_cells * sizeof(double));
  out->volume = (double *)malloc(6 * solid_cells * sizeof(double));
  out->lane = (uint32_t *)malloc(6 * solid_cells * sizeof(uint32_t));
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
  /* anchored: inside a support volume (solver.cc:163-180, 2265-2268) */
  for (uint32_t n = 0; n < nodes; ++n) {
    for (uint32_t v = 0; v < nsupports; ++v) {
      const float *M = supports[v].world_to_support;
      int inside = 1;
      for (int r = 0; r < 3 && inside; ++r) {
        const float l = M[r * 4 + 0] * restf[3 * n + 0] +
                        M[r * 4 + 1] * restf[3 * n + 1] +
                        M[r * 4 + 2] * restf[3 * n + 2] + M[r * 4 + 3];
        if (!(fabsf(l) <= supports[v].half[r])) inside = 0;
      }
      if (inside) {
        out->anchored[n] = 1;
        break;
      }
    }
  }
  uint32_t tets = 0;
  orc_material fb;
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        const size_t cell = x + (size_t)cx * (y + (size_t)cy * z);
        if (L->solid && !L->solid[cell]) continue;
        uint32_t corners[8];
        for (uint32_t c = 0; c < 8; ++c) {
          const uint32_t ax = x + (c & 1u), ay = y + ((c >> 1) & 1u),
                         az = z + ((c >> 2) & 1u);
          corners[c] =
              node_of[ax + (size_t)L->res[0] * (ay + (size_t)L->res[1] * az)];
        }
        const uint32_t lane = L->cell_lane ? L->cell_lane[cell] : 0u;
        const double density =
            (double)lane_of(materials, lanes, lane, &fb)->density;
        for (int k = 0; k < 6; ++k) {
          uint32_t tn[4] = {corners[KUHN[k][0]], corners[KUHN[k][1]],
                            corners[KUHN[k][2]], corners[KUHN[k][3]]};
          double shape[9];
          for (int pass = 0; pass < 2; ++pass) {
            for (int e = 0; e < 3; ++e)
              for (int r = 0; r < 3; ++r)
                shape[r * 3 + e] =
                    out->rest[3 * tn[e + 1] + r] - out->rest[3 * tn[0] + r];
            if (pass == 0 && orc_det3(shape) < 0.0) {
              const uint32_t t = tn[1];
              tn[1] = tn[2];
              tn[2] = t;
              continue;
            }
            break;
          }
          const double vol = orc_det3(shape) / 6.0;
          if (!(vol > 0.0)) continue;
          memcpy(&out->tet[4 * (size_t)tets], tn, sizeof tn);
          inv3(shape, &out->dminv[9 * (size_t)tets]);
          out->volume[tets] = vol;
          out->lane[tets] = lane;
          for (int c = 0; c < 4; ++c) out->mass[tn[c]] += density * vol * 0.25;
          ++tets;
        }
      }
  out->tets = tets;
  free(restf);
  free(node_of);
  return tets ? 0 : -1;
}

void orc_mesh_free(orc_mesh *m) {
  free(m->rest);
  free(m->mass);
  free(m->anchored);
  free(m->tet);
  free(m->dminv);
  free(m->volume);
  free(m->lane);
  memset(m, 0, sizeof *m);
}

/* ------------------------------------------------------------------ */
/* Rotations                                                            */
/* ------------------------------------------------------------------ */

/* polarRotation (material.cc:168-197): R <- (R + R^-T)/2 from R = F, at most
 * 24 passes, stop when the largest entry change is below 1e-6; identity when
 * |det F| < 1e-12. */
void orc_polar(const double *F, double *R) {
  if (fabs(orc_det3(F)) < 1.0e-12) {
    ident3(R);
    return;
  }
  double r[9];
  memcpy(r, F, sizeof r);
  for (int pass = 0; pass < 24; ++pass) {
    double c[9];
    cof3(r, c);
    const double d = r[0] * c[0] + r[1] * c[1] + r[2] * c[2];
    double largest = 0.0;
    for (int i = 0; i < 9; ++i) {
      const double next = 0.5 * (r[i] + c[i] / d); /* R^-T = cof(R)/det */
      const double change = fabs(next - r[i]);
      if (change > largest) largest = change;
      r[i] = next;
    }
    if (largest < 1.0e-6) break;
  }
  memcpy(R, r, sizeof r);
}

/* elementRotation, Invertible mode (material.cc:24-64): polar when det F > 0,
 * else the rotation of a sign-fixed SVD, U diag(1,1,det(UV^T)) V^T with the
 * smallest singular value carrying the reflection.  Computed from the
 * eigenvectors of F^T F; unique whenever sigma_2 > 0. */
void orc_element_rotation(const double *F, double *R) {
  if (orc_det3(F) > 0.0) {
    orc_polar(F, R);
    return;
  }
  double C[9], w[3], V[9];
  mul3_tn(F, F, C);
  orc_eigen_sym3(C, w, V); /* ascending: column 2 = largest */
  /* v1 = largest, v2 = middle, v3 = smallest; make (v1,v2,v3) proper */
  double v[3][3];
  for (int r = 0; r < 3; ++r) {
    v[0][r] = V[r * 3 + 2];
    v[1][r] = V[r * 3 + 1];
    v[2][r] = V[r * 3 + 0];
  }
  double cr[3] = {v[0][1] * v[1][2] - v[0][2] * v[1][1],
                  v[0][2] * v[1][0] - v[0][0] * v[1][2],
                  v[0][0] * v[1][1] - v[0][1] * v[1][0]};
  for (int r = 0; r < 3; ++r) v[2][r] = cr[r];
  double u[3][3];
  for (int i = 0; i < 2; ++i) mv3(F, v[i], u[i]);
  double n0 = sqrt(u[0][0] * u[0][0] + u[0][1] * u[0][1] + u[0][2] * u[0][2]);
  if (!(n0 > 0.0)) {
    ident3(R);
    return;
  }
  for (int r = 0; r < 3; ++r) u[0][r] /= n0;
  const double p = u[1][0] * u[0][0] + u[1][1] * u[0][1] + u[1][2] * u[0][2];
  for (int r = 0; r < 3; ++r) u[1][r] -= p * u[0][r];
  double n1 = sqrt(u[1][0] * u[1][0] + u[1][1] * u[1][1] + u[1][2] * u[1][2]);
  if (!(n1 > 0.0)) {
    ident3(R);
    return;
  }
  for (int r = 0; r < 3; ++r) u[1][r] /= n1;
  u[2][0] = u[0][1] * u[1][2] - u[0][2] * u[1][1];
  u[2][1] = u[0][2] * u[1][0] - u[0][0] * u[1][2];
  u[2][2] = u[0][0] * u[1][1] - u[0][1] * u[1][0];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      R[r * 3 + c] = u[0][r] * v[0][c] + u[1][r] * v[1][c] + u[2][r] * v[2][c];
}

/* ------------------------------------------------------------------ */
/* Stable neo-Hookean, one point per tet, displacement (H) form         */
/* ------------------------------------------------------------------ */

/* SNH energy density and first Piola stress from H = F - I
 * (neohookean.cc:41-47, 125-147; material.cc:918-946):
 *   J-1 = tr H + (tr^2 H - H:H^T)/2 + det H
 *   psi = mu/2 |H|^2 - mu (J-1 - tr H) + (lambda+mu)/2 (J-1)^2
 *   P   = mu (H + H^T - tr H I - cof H) + (lambda+mu)(J-1) cof F        */
static double snh_stress(const double *H, double lambda, double mu,
                         double *P) {
  const double shifted = lambda + mu;
  const double tr = trace3(H);
  double Ht_contract = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Ht_contract += H[r * 3 + c] * H[c * 3 + r];
  const double change = tr + 0.5 * (tr * tr - Ht_contract) + orc_det3(H);
  double cofH[9], F[9], cofF[9];
  cof3(H, cofH);
  for (int i = 0; i < 9; ++i) F[i] = H[i];
  F[0] += 1.0;
  F[4] += 1.0;
  F[8] += 1.0;
  cof3(F, cofF);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      P[r * 3 + c] = mu * (H[r * 3 + c] + H[c * 3 + r] - (r == c ? tr : 0.0) -
                           cofH[r * 3 + c]) +
                     shifted * change * cofF[r * 3 + c];
  return 0.5 * mu * contract3(H, H) - mu * (change - tr) +
         0.5 * shifted * change * change;
}

/* Forces f_i = -V P g_i, or with a plastic offset Ep = Fp^-1 - I the elastic
 * gradient F_e = F (I + Ep) and f_i = -V P(F_e) (I+Ep)^T g_i. */
void orc_snh_tet(const double *u, const double *dminv, double volume,
                 double lambda, double mu, const double *ep, double *force,
                 double *energy) {
  double H[9], g[4][3], P[9];
  displacement_gradient(u, dminv, H);
  shape_gradients(dminv, g);
  double He[9];
  if (ep) {
    double HE[9];
    mul3(H, ep, HE);
    for (int i = 0; i < 9; ++i) He[i] = H[i] + ep[i] + HE[i];
  } else {
    memcpy(He, H, sizeof He);
  }
  const double psi = snh_stress(He, lambda, mu, P);
  double PF[9];
  if (ep) {
    double Fp[9]; /* (I + Ep) */
    memcpy(Fp, ep, sizeof Fp);
    Fp[0] += 1.0;
    Fp[4] += 1.0;
    Fp[8] += 1.0;
    mul3_nt(P, Fp, PF); /* P (I+Ep)^T */
  } else {
    memcpy(PF, P, sizeof PF);
  }
  for (int i = 0; i < 4; ++i) {
    double t[3];
    mv3(PF, g[i], t);
    for (int r = 0; r < 3; ++r) force[3 * i + r] = -volume * t[r];
  }
  if (energy) *energy = volume * psi;
}

/* ------------------------------------------------------------------ */
/* Corotated tet with optional J2 return map                            */
/* ------------------------------------------------------------------ */

/* Linearisation of one return map (PointResponse, material.cc:581-592):
 * theta, thetaBar and the flow normal of a loading point; the default
 * (theta 1, thetaBar 0) is the elastic law. */
typedef struct {
  double theta, theta_bar, normal[9], excess;
  int plastic;
} j2_lin_t;

/* returnMap (material.cc:594-630) on the small strain in the corotated frame,
 * plastic = {eps_p (9, row-major), alpha}.  Returns 1 when the point returns
 * to the yield surface; return_mapping = 0 is the elastic trial
 * (returnMapping false).  lin, if given, receives the linearisation. */
static int j2_return_map(const double *strain, const double *plastic, double lambda, double mu,
                         double yield, double hardening, int return_mapping, double *stress,
                         double *next, j2_lin_t *lin) {
  const double bulk = lambda + 2.0 * mu / 3.0;
  double el[9];
  for (int i = 0; i < 9; ++i) el[i] = strain[i] - (plastic ? plastic[i] : 0.0);
  const double tr = trace3(el);
  const double pressure = bulk * tr;
  double trial[9];
  for (int i = 0; i < 9; ++i) trial[i] = 2.0 * mu * el[i];
  for (int d = 0; d < 3; ++d) trial[d * 4] -= 2.0 * mu * tr / 3.0;
  const double norm = sqrt(contract3(trial, trial));
  const double equivalent = sqrt(1.5) * norm;
  const double alpha = plastic ? plastic[9] : 0.0;
  const double radius = yield + hardening * alpha;
  if (next && plastic) memcpy(next, plastic, 10 * sizeof(double));
  if (next && !plastic) memset(next, 0, 10 * sizeof(double));
  if (lin) {
    memset(lin, 0, sizeof *lin);
    lin->theta = 1.0;
  }
  if (!return_mapping || !(equivalent > radius) || !(norm > 0.0)) {
    for (int i = 0; i < 9; ++i) stress[i] = trial[i];
    for (int d = 0; d < 3; ++d) stress[d * 4] += pressure;
    return 0;
  }
  const double gamma = (equivalent - radius) / (3.0 * mu + hardening);
  const double theta = 1.0 - 3.0 * mu * gamma / equivalent;
  for (int i = 0; i < 9; ++i) stress[i] = theta * trial[i];
  for (int d = 0; d < 3; ++d) stress[d * 4] += pressure;
  if (next) {
    for (int i = 0; i < 9; ++i) next[i] += sqrt(1.5) * gamma * trial[i] / norm;
    next[9] += gamma;
  }
  if (lin) {
    lin->theta = theta;
    lin->theta_bar = 3.0 * mu / (3.0 * mu + hardening) - (1.0 - theta);
    for (int i = 0; i < 9; ++i) lin->normal[i] = trial[i] / norm;
    lin->excess = (equivalent - radius) / radius;
    lin->plastic = 1;
  }
  return 1;
}

/* tangentOf (material.cc:636-647): the consistent tangent's stress for a
 * strain e, K tr(e) I + 2 mu theta dev(e) - 2 mu thetaBar (n:e) n. */
static void j2_tangent_stress(const j2_lin_t *lin, double lambda, double mu, const double *e,
                              double *out) {
  const double bulk = lambda + 2.0 * mu / 3.0;
  const double tr = trace3(e);
  const double along = contract3(lin->normal, e);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      out[r * 3 + c] = 2.0 * mu * lin->theta * (e[r * 3 + c] - (r == c ? tr / 3.0 : 0.0)) +
                       (r == c ? bulk * tr : 0.0) -
                       2.0 * mu * lin->theta_bar * along * lin->normal[r * 3 + c];
}

/* tetStrain + trial / tetPointTrial (material.cc:136-164, 236-265,
 * 978-1031): R = elementRotation(F); centred offsets
 * o_i = R^T (x_i - xbar) - (X_i - Xbar); eps = sym(sum o_i g_i^T);
 * sigma from the return map (elastic when the lane does not yield or
 * return_mapping is 0); f_i = -R V sigma g_i; energy V/2 sigma:(eps -
 * eps_p_next).  lin receives the return map's linearisation. */
static void corotated_eval(const double *x, const double *X, const double *dminv, double volume,
                           double lambda, double mu, const double *plastic, double yield,
                           double hardening, int return_mapping, double *R, double *force,
                           double *energy, double *plastic_next, uint32_t *yielding, j2_lin_t *lin) {
  double g[4][3], F[9], u[12];
  shape_gradients(dminv, g);
  for (int i = 0; i < 12; ++i) u[i] = x[i] - X[i];
  displacement_gradient(u, dminv, F);
  F[0] += 1.0;
  F[4] += 1.0;
  F[8] += 1.0;
  orc_element_rotation(F, R);
  double xc[3] = {0, 0, 0}, Xc[3] = {0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int r = 0; r < 3; ++r) {
      xc[r] += 0.25 * x[3 * i + r];
      Xc[r] += 0.25 * X[3 * i + r];
    }
  double D[9] = {0};
  for (int i = 0; i < 4; ++i) {
    double arm[3], o[3];
    for (int r = 0; r < 3; ++r) arm[r] = x[3 * i + r] - xc[r];
    mtv3(R, arm, o);
    for (int r = 0; r < 3; ++r) o[r] -= X[3 * i + r] - Xc[r];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) D[r * 3 + c] += o[r] * g[i][c];
  }
  double eps[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) eps[r * 3 + c] = 0.5 * (D[r * 3 + c] + D[c * 3 + r]);
  double sigma[9], next[10];
  const int yields_now = isfinite(yield) && yield > 0.0;
  const int plastic_now =
      yields_now ? j2_return_map(eps, plastic, lambda, mu, yield, hardening, return_mapping, sigma, next, lin)
                 : j2_return_map(eps, NULL, lambda, mu, INFINITY, 0.0, 1, sigma, next, lin);
  if (!yields_now && plastic) memcpy(next, plastic, 10 * sizeof(double));
  for (int i = 0; i < 4; ++i) {
    double t[3], f[3];
    mv3(sigma, g[i], t);
    mv3(R, t, f);
    for (int r = 0; r < 3; ++r) force[3 * i + r] = -volume * f[r];
  }
  double el[9];
  for (int i = 0; i < 9; ++i) el[i] = eps[i] - next[i];
  if (energy) *energy = volume * 0.5 * contract3(sigma, el);
  if (plastic_next) memcpy(plastic_next, next, 10 * sizeof(double));
  if (yielding) *yielding = (uint32_t)plastic_now;
}

void orc_corotated_tet(const double *x, const double *X, const double *dminv,
                       double volume, double lambda, double mu,
                       const double *plastic, double yield, double hardening,
                       double *R, double *force, double *energy,
                       double *plastic_next, uint32_t *yielding) {
  corotated_eval(x, X, dminv, volume, lambda, mu, plastic, yield, hardening, 1, R, force, energy,
                 plastic_next, yielding, NULL);
}

/* ------------------------------------------------------------------ */
/* Hencky return and plastic flow for SNH tets (plasticflow.cc:96-260)  */
/* ------------------------------------------------------------------ */

/* returnMap (plasticflow.cc:110-150): log principal strains of F_e F_e^T,
 * radial return of the deviator to the yield surface.  Returns 1 and the
 * returned F_e when the trial stress exceeds the yield stress. */
static int hencky_return(const double *elastic, double mu, double yield,
                         double *returned, double *trial_stress) {
  double b[9], w[3], V[9];
  mul3_nt(elastic, elastic, b);
  orc_eigen_sym3(b, w, V);
  if (!(w[0] > 0.0)) return 0;
  double strain[3], dev[3];
  for (int i = 0; i < 3; ++i) strain[i] = 0.5 * log(w[i]);
  const double mean = (strain[0] + strain[1] + strain[2]) / 3.0;
  for (int i = 0; i < 3; ++i) dev[i] = strain[i] - mean;
  const double size = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const double stress = sqrt(1.5) * 2.0 * mu * size;
  if (trial_stress) *trial_stress = stress;
  if (!(stress > yield) || size <= 1.0e-15) return 0;
  double scale[3];
  for (int i = 0; i < 3; ++i) scale[i] = exp(dev[i] * (yield / stress - 1.0));
  double corr[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s += V[r * 3 + i] * scale[i] * V[c * 3 + i];
      corr[r * 3 + c] = s;
    }
  mul3(corr, elastic, returned);
  return 1;
}

/* flowLattice body for one point (plasticflow.cc:170-255), applied to the
 * single point of a tet.  plastic_offset = Fp^-1 - I (updated in place),
 * alpha = accumulated hardening strain.  Returns 1 on a flow (yield) event.
 * Crush compaction (crushStress) is not restated. */
int orc_hencky_flow(const double *F, double mu, double yield, double hardening,
                    double plastic_max, double *ep, double *alpha) {
  if (!(orc_det3(F) > 0.0)) return 0;
  double plastic[9], elastic[9];
  memcpy(plastic, ep, sizeof plastic);
  plastic[0] += 1.0;
  plastic[4] += 1.0;
  plastic[8] += 1.0;
  mul3(F, plastic, elastic);
  const double start = yield + hardening * (*alpha);
  const double ratio = hardening / (3.0 * mu);
  double after[9], sigma_trial = 0.0, returned[9];
  const int trial = hencky_return(elastic, mu, start, returned, &sigma_trial);
  if (!trial) return 0;
  const double yield_now = (start + ratio * sigma_trial) / (1.0 + ratio);
  double s2 = 0.0;
  if (!hencky_return(elastic, mu, yield_now, after, &s2)) return 0;
  double Finv[9], next[9];
  if (!inv3(F, Finv)) return 0;
  mul3(Finv, after, next);
  const double kept = orc_det3(next);
  if (isfinite(plastic_max) && kept > 0.0) {
    double ninv[9];
    inv3(next, ninv);
    ninv[0] -= 1.0;
    ninv[4] -= 1.0;
    ninv[8] -= 1.0;
    const double size = sqrt(contract3(ninv, ninv));
    if (size > plastic_max) {
      double capped[9], cinv[9];
      for (int i = 0; i < 9; ++i) capped[i] = (plastic_max / size) * ninv[i];
      capped[0] += 1.0;
      capped[4] += 1.0;
      capped[8] += 1.0;
      const double cv = orc_det3(capped);
      if (!(cv > 1.0e-6)) return 0;
      inv3(capped, cinv);
      const double s = cbrt(kept * cv);
      for (int i = 0; i < 9; ++i) next[i] = s * cinv[i];
    }
  }
  double change = 0.0;
  for (int i = 0; i < 9; ++i) change += fabs(next[i] - plastic[i]);
  if (!(change > 1.0e-12)) return 0;
  for (int i = 0; i < 9; ++i) ep[i] = next[i] - (i % 4 == 0 ? 1.0 : 0.0);
  *alpha += (sigma_trial - yield_now) / (3.0 * mu);
  return 1;
}

/* stretchOf (tearing.cc:248-290) for a tet: the edge frame at every corner
 * of a tet gives the same F, so s = sqrt(lambda_max(F^T F)); direction is the
 * top eigenvector (material frame). */
double orc_stretch(const double *F, double *direction) {
  double C[9], w[3], V[9];
  mul3_tn(F, F, C);
  orc_eigen_sym3(C, w, V);
  if (direction) {
    double n = 0.0;
    for (int r = 0; r < 3; ++r) n += V[r * 3 + 2] * V[r * 3 + 2];
    n = sqrt(n);
    for (int r = 0; r < 3; ++r) direction[r] = V[r * 3 + 2] / n;
  }
  return sqrt(w[2] > 0.0 ? w[2] : 0.0);
}

/* ------------------------------------------------------------------ */
/* Threads: contiguous chunks; every output has one writer              */
/* ------------------------------------------------------------------ */

typedef void (*range_fn)(void *ctx, size_t begin, size_t end);
typedef struct {
  range_fn fn;
  void *ctx;
  size_t begin, end;
} range_job;

static void *range_main(void *p) {
  range_job *j = (range_job *)p;
  j->fn(j->ctx, j->begin, j->end);
  return NULL;
}

static void parallel_for(size_t n, uint32_t threads, range_fn fn, void *ctx) {
  if (threads <= 1 || n < 4096) {
    fn(ctx, 0, n);
    return;
  }
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  range_job job[256];
  const size_t chunk = (n + threads - 1) / threads;
  uint32_t started = 0;
  for (uint32_t t = 0; t < threads; ++t) {
    const size_t b = (size_t)t * chunk;
    if (b >= n) break;
    job[t].fn = fn;
    job[t].ctx = ctx;
    job[t].begin = b;
    job[t].end = b + chunk < n ? b + chunk : n;
    if (t == 0) continue;
    pthread_create(&tid[t], NULL, range_main, &job[t]);
    ++started;
  }
  fn(ctx, job[0].begin, job[0].end);
  for (uint32_t t = 1; t <= started; ++t) pthread_join(tid[t], NULL);
}

/* deterministic dot over n vec3: fixed 1024-node blocks summed in order */
#define DOT_BLOCK 1024u
typedef struct {
  const double *a, *b;
  const uint8_t *mask;
  size_t n;
  double *partial;
} dot_ctx;

static void dot_range(void *p, size_t begin, size_t end) {
  dot_ctx *c = (dot_ctx *)p;
  for (size_t blk = begin; blk < end; ++blk) {
    double s = 0.0;
    const size_t lo = blk * DOT_BLOCK;
    const size_t hi = lo + DOT_BLOCK < c->n ? lo + DOT_BLOCK : c->n;
    for (size_t i = lo; i < hi; ++i)
      s += c->a[3 * i] * c->b[3 * i] + c->a[3 * i + 1] * c->b[3 * i + 1] +
           c->a[3 * i + 2] * c->b[3 * i + 2];
    c->partial[blk] = s;
  }
}

/* ------------------------------------------------------------------ */
/* Bodies                                                               */
/* ------------------------------------------------------------------ */

struct orc_body {
  orc_mesh m;
  double *lambda, *mu, *yield, *hard, *pmax, *tear, *rho;
  uint8_t *snh, *plastic_lane;
  double *x, *v, *a;
  double *R;       /* 9 per tet, committed rotation */
  double *W;       /* 9 per tet, neo-Hookean volume frame cof(Fe) Fp^-T */
  double *plastic; /* 10 per tet */
  uint8_t *yflag;
  /* plastic Newton of yielding corotated tets: the assembly's return-map
   * switch, loading margin and stiffness binding (bindStiffness: the
   * consistent tangent when set), the per-tet linearisation (theta,
   * thetaBar, normal[9]) and loading flags */
  int newton, tangent, return_mapping;
  double margin;
  double *lin;
  uint8_t *loading;
  double *total, *lowest_it;
  /* stiffness damping (factor.cc:755-855, solver.cc:625-902): per tet tau
   * and rigid fit (arms 12, inverse inertia 9) at the last assembly, the
   * operator's damping coefficient (1 - alpha_f) gamma h, per-slot damping
   * values and diagonal blocks */
  int damped;
  double *tau, *fit, *slot_v, *slot_dd;
  double damp, lag;
  uint32_t *inc_start, *inc_slot;
  /* scratch */
  double *slot_f, *slot_d, *tet_energy;
  double *force, *diag, *precond, *right, *weighted, *shifted, *carried,
      *drive, *evaluated, *residual, *direction, *applied, *conditioned,
      *increment, *unconstrained;
  double *partial;
  size_t nblocks;
  /* rollout checkpoints (x, v, a every `every` steps) and final state */
  double *ckpt, *final_state;
  uint32_t ckpt_steps, ckpt_every;
  orc_settings ckpt_settings;
};

orc_body *orc_body_create(const orc_mesh *mesh, const orc_material *materials,
                          uint32_t lanes) {
  orc_body *b = (orc_body *)calloc(1, sizeof *b);
  const size_t N = mesh->nodes, T = mesh->tets;
  b->m.nodes = mesh->nodes;
  b->m.tets = mesh->tets;
#define DUP(dst, src, count, type)                        \
  do {                                                    \
    dst = (type *)malloc((count) * sizeof(type) + 1);     \
    memcpy(dst, src, (count) * sizeof(type));             \
  } while (0)
  DUP(b->m.rest, mesh->rest, 3 * N, double);
  DUP(b->m.mass, mesh->mass, N, double);
  DUP(b->m.anchored, mesh->anchored, N, uint8_t);
  DUP(b->m.tet, mesh->tet, 4 * T, uint32_t);
  DUP(b->m.dminv, mesh->dminv, 9 * T, double);
  DUP(b->m.volume, mesh->volume, T, double);
  DUP(b->m.lane, mesh->lane, T, uint32_t);
#undef DUP
  b->lambda = (double *)malloc(T * sizeof(double));
  b->mu = (double *)malloc(T * sizeof(double));
  b->yield = (double *)malloc(T * sizeof(double));
  b->hard = (double *)malloc(T * sizeof(double));
  b->pmax = (double *)malloc(T * sizeof(double));
  b->tear = (double *)malloc(T * sizeof(double));
  b->rho = (double *)malloc(T * sizeof(double) + 8);
  b->snh = (uint8_t *)malloc(T + 1);
  b->plastic_lane = (uint8_t *)malloc(T + 1);
  orc_material fb;
  for (size_t e = 0; e < T; ++e) {
    const orc_material *mt = lane_of(materials, lanes, mesh->lane[e], &fb);
    orc_lame(mt->youngs_modulus, mt->poissons_ratio, &b->lambda[e], &b->mu[e]);
    b->plastic_lane[e] = (uint8_t)yields(mt->yield_stress);
    b->yield[e] = b->plastic_lane[e] ? (double)mt->yield_stress : INFINITY;
    b->hard[e] = mt->hardening_modulus > 0.0f ? (double)mt->hardening_modulus : 0.0;
    b->pmax[e] = (double)mt->plastic_max;
    b->tear[e] = 1.0 + (double)mt->tear_strain;
    b->rho[e] = (double)mt->density;
    b->snh[e] = mt->elastic_model == 1u ? 1u : 0u;
  }
  b->x = (double *)malloc(3 * N * sizeof(double));
  memcpy(b->x, mesh->rest, 3 * N * sizeof(double));
  b->v = (double *)calloc(3 * N, sizeof(double));
  b->a = (double *)calloc(3 * N, sizeof(double));
  b->R = (double *)malloc(9 * T * sizeof(double));
  b->W = (double *)malloc(9 * T * sizeof(double));
  for (size_t e = 0; e < T; ++e) {
    ident3(&b->R[9 * e]);
    ident3(&b->W[9 * e]);
  }
  b->plastic = (double *)calloc(10 * T, sizeof(double));
  b->yflag = (uint8_t *)calloc(T + 1, 1);
  b->lin = (double *)calloc(11 * T + 1, sizeof(double));
  b->loading = (uint8_t *)calloc(T + 1, 1);
  b->total = (double *)calloc(3 * N + 1, sizeof(double));
  b->lowest_it = (double *)calloc(3 * N + 1, sizeof(double));
  b->return_mapping = 1;
  b->tau = (double *)calloc(T + 1, sizeof(double));
  for (size_t e = 0; e < T; ++e) {
    const orc_material *mt = lane_of(materials, lanes, mesh->lane[e], &fb);
    const double sd = (double)mt->stiffness_damping;
    b->tau[e] = isfinite(sd) && sd > 0.0 ? sd : 0.0; /* stiffnessDamping, factor.cc:755-759 */
    if (b->tau[e] > 0.0) b->damped = 1;
  }
  b->fit = (double *)calloc(21 * T + 1, sizeof(double));
  b->slot_v = (double *)calloc(12 * T + 1, sizeof(double));
  b->slot_dd = (double *)calloc(36 * T + 1, sizeof(double));
  for (size_t e = 0; e < T; ++e) {
    b->lin[11 * e] = 1.0;
    if (b->plastic_lane[e] && !b->snh[e]) b->newton = 1;
  }
  /* incidence: node -> slots (tet*4+corner) in ascending slot order
   * (buildIncidence, solver.cc:1065-1101) */
  b->inc_start = (uint32_t *)calloc(N + 1, sizeof(uint32_t));
  for (size_t s = 0; s < 4 * T; ++s) ++b->inc_start[mesh->tet[s] + 1];
  for (size_t n = 0; n < N; ++n) b->inc_start[n + 1] += b->inc_start[n];
  b->inc_slot = (uint32_t *)malloc(4 * T * sizeof(uint32_t) + 4);
  uint32_t *cursor = (uint32_t *)malloc(N * sizeof(uint32_t) + 4);
  memcpy(cursor, b->inc_start, N * sizeof(uint32_t));
  for (size_t s = 0; s < 4 * T; ++s) b->inc_slot[cursor[mesh->tet[s]]++] = (uint32_t)s;
  free(cursor);
  b->slot_f = (double *)malloc(12 * T * sizeof(double));
  b->slot_d = (double *)malloc(36 * T * sizeof(double));
  b->tet_energy = (double *)malloc(T * sizeof(double) + 8);
#define VEC(p) p = (double *)calloc(3 * N, sizeof(double))
  VEC(b->force);
  VEC(b->right);
  VEC(b->weighted);
  VEC(b->shifted);
  VEC(b->carried);
  VEC(b->drive);
  VEC(b->evaluated);
  VEC(b->residual);
  VEC(b->direction);
  VEC(b->applied);
  VEC(b->conditioned);
  VEC(b->increment);
  VEC(b->unconstrained);
#undef VEC
  b->diag = (double *)calloc(9 * N, sizeof(double));
  b->precond = (double *)calloc(9 * N, sizeof(double));
  b->nblocks = (N + DOT_BLOCK - 1) / DOT_BLOCK;
  b->partial = (double *)calloc(b->nblocks + 1, sizeof(double));
  return b;
}

void orc_body_free(orc_body *b) {
  if (!b) return;
  orc_mesh_free(&b->m);
  void *ptrs[] = {b->lambda, b->mu, b->yield, b->hard, b->pmax, b->tear, b->rho,
                  b->snh, b->plastic_lane, b->x, b->v, b->a, b->R, b->W, b->plastic,
                  b->yflag, b->inc_start, b->inc_slot, b->slot_f, b->slot_d,
                  b->tet_energy, b->force, b->diag, b->precond, b->right,
                  b->weighted, b->shifted, b->carried, b->drive, b->evaluated,
                  b->residual, b->direction, b->applied, b->conditioned,
                  b->increment, b->unconstrained, b->partial, b->ckpt, b->final_state,
                  b->lin, b->loading, b->total, b->lowest_it, b->tau, b->fit, b->slot_v, b->slot_dd};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; ++i) free(ptrs[i]);
  free(b);
}

void orc_body_set_state(orc_body *b, const double *x, const double *v,
                        const double *a) {
  const size_t n = 3 * (size_t)b->m.nodes * sizeof(double);
  if (x) memcpy(b->x, x, n);
  if (v) memcpy(b->v, v, n);
  if (a) memcpy(b->a, a, n);
}

void orc_body_get_state(const orc_body *b, double *x, double *v, double *a) {
  const size_t n = 3 * (size_t)b->m.nodes * sizeof(double);
  if (x) memcpy(x, b->x, n);
  if (v) memcpy(v, b->v, n);
  if (a) memcpy(a, b->a, n);
}

void orc_body_get_rotations(const orc_body *b, double *R) {
  memcpy(R, b->R, 9 * (size_t)b->m.tets * sizeof(double));
}
void orc_body_get_frames(const orc_body *b, double *W) {
  memcpy(W, b->W, 9 * (size_t)b->m.tets * sizeof(double));
}

void orc_body_set_frames(orc_body *b, const double *W) {
  memcpy(b->W, W, 9 * (size_t)b->m.tets * sizeof(double));
}

void orc_body_set_rotations(orc_body *b, const double *R) {
  memcpy(b->R, R, 9 * (size_t)b->m.tets * sizeof(double));
}
void orc_body_get_plastic(const orc_body *b, double *p) {
  memcpy(p, b->plastic, 10 * (size_t)b->m.tets * sizeof(double));
}
void orc_body_set_plastic(orc_body *b, const double *p) {
  memcpy(b->plastic, p, 10 * (size_t)b->m.tets * sizeof(double));
}
void orc_yield_flags(const orc_body *b, uint8_t *flags) {
  memcpy(flags, b->yflag, b->m.tets);
}

/* ---- element pass ---- */

typedef struct {
  orc_body *b;
  const double *at;  /* positions */
  int diag;          /* write block-Jacobi diagonal slots */
} elem_ctx;

static void gather4(const orc_body *b, const double *src, size_t e,
                    double *out) {
  for (int i = 0; i < 4; ++i)
    for (int r = 0; r < 3; ++r)
      out[3 * i + r] = src[3 * (size_t)b->m.tet[4 * e + i] + r];
}

/* Neo-Hookean volume frame W = cof(Fe) Fp^-T with Fe = F Fp^-1 and
 * Fp^-1 = I + ep: the derivative of det(Fe) along dF is W : dF. */
static void volume_frame(const double *F, const double *ep, double *W) {
  if (!ep) {
    cof3(F, W);
    return;
  }
  double Fi[9], Fe[9], C[9];
  memcpy(Fi, ep, sizeof Fi);
  Fi[0] += 1.0;
  Fi[4] += 1.0;
  Fi[8] += 1.0;
  mul3(F, Fi, Fe);
  cof3(Fe, C);
  mul3_nt(C, Fi, W);
}

/* ---- stiffness damping (factor.cc:755-855; solver.cc:625-902) ---- */

/* pseudoInverse (factor.cc:183-272) of a symmetric 3x3: eigenvalues above
 * `floor` inverted, the others dropped */
static void pseudo_inverse_sym3(const double *A, double floor, double *out) {
  double w[3], V[9];
  orc_eigen_sym3(A, w, V);
  memset(out, 0, 9 * sizeof(double));
  for (int k = 0; k < 3; ++k) {
    if (!(w[k] > floor)) continue;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) out[r * 3 + c] += V[r * 3 + k] * V[c * 3 + k] / w[k];
  }
}

/* rigidFit (factor.cc:783-817) of a tet's corners: arms from the centre and
 * the inverse of sum (|a|^2 I - a a^T), a pseudo-inverse when nearly
 * singular.  fit = {arms[12], inverse[9]} */
static void rigid_fit(const double *corners, double *fit) {
  double c[3] = {0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int r = 0; r < 3; ++r) c[r] += corners[3 * i + r] / 4.0;
  double I[9] = {0};
  for (int i = 0; i < 4; ++i) {
    double *a = &fit[3 * i];
    for (int r = 0; r < 3; ++r) a[r] = corners[3 * i + r] - c[r];
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) I[r * 3 + k] += (r == k ? aa : 0.0) - a[r] * a[k];
  }
  const double third = (I[0] + I[4] + I[8]) / 3.0;
  if (!(orc_det3(I) > 1.0e-6 * third * third * third) || !inv3(I, &fit[12]))
    pseudo_inverse_sym3(I, 1.0e-6 * third, &fit[12]);
}

/* removeRigid (factor.cc:819-846): v_i -= mean + spin x arm_i,
 * spin = inverse * sum arm_i x v_i */
static void remove_rigid(const double *fit, double *v) {
  double mean[3] = {0, 0, 0}, moment[3] = {0, 0, 0}, spin[3];
  for (int i = 0; i < 4; ++i) {
    const double *a = &fit[3 * i], *x = &v[3 * i];
    for (int r = 0; r < 3; ++r) mean[r] += x[r] / 4.0;
    moment[0] += a[1] * x[2] - a[2] * x[1];
    moment[1] += a[2] * x[0] - a[0] * x[2];
    moment[2] += a[0] * x[1] - a[1] * x[0];
  }
  mv3(&fit[12], moment, spin);
  for (int i = 0; i < 4; ++i) {
    const double *a = &fit[3 * i];
    double *x = &v[3 * i];
    const double sa[3] = {spin[1] * a[2] - spin[2] * a[1], spin[2] * a[0] - spin[0] * a[2],
                          spin[0] * a[1] - spin[1] * a[0]};
    for (int r = 0; r < 3; ++r) x[r] -= mean[r] + sa[r];
  }
}

/* applyDamping (factor.cc:848-857): out = P R K R^T P v with P the rigid-
 * mode removal of the fit and K the tet's elastic stiffness */
static void damp_tet(const orc_body *b, size_t e, const double *v, double *out) {
  const double *fit = &b->fit[21 * e], *R = &b->R[9 * e];
  double rel[12], g[4][3], D[9] = {0};
  memcpy(rel, v, sizeof rel);
  remove_rigid(fit, rel);
  shape_gradients(&b->m.dminv[9 * e], g);
  for (int j = 0; j < 4; ++j) {
    double dp[3];
    mtv3(R, &rel[3 * j], dp);
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) D[r * 3 + k] += dp[r] * g[j][k];
  }
  const double tr = trace3(D), V = b->m.volume[e];
  double sg[9];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      sg[r * 3 + k] = b->mu[e] * (D[r * 3 + k] + D[k * 3 + r]) + (r == k ? b->lambda[e] * tr : 0.0);
  for (int i = 0; i < 4; ++i) {
    double t[3];
    mv3(sg, g[i], t);
    mv3(R, t, &out[3 * i]);
    for (int r = 0; r < 3; ++r) out[3 * i + r] *= V;
  }
  remove_rigid(fit, out);
}

/* Shape gradients of the plastic reference, Fp^-T g = (I + ep)^T g, for a
 * plastic neo-Hookean tet; the rest gradients otherwise. */
static void plastic_gradients(const orc_body *b, size_t e, double g[4][3], double gs[4][3]) {
  const int plastic = b->snh[e] && b->plastic_lane[e];
  const double *ep = &b->plastic[10 * e];
  for (int i = 0; i < 4; ++i)
    for (int r = 0; r < 3; ++r) {
      double s = g[i][r];
      if (plastic)
        for (int k = 0; k < 3; ++k) s += ep[k * 3 + r] * g[i][k];
      gs[i][r] = s;
    }
}

/* One tet of assembleElastic (solver.cc:1382-1470): committed rotation,
 * forces to slots, diagonal stiffness blocks to slots.  Corotated:
 * R K_ii R^T = V((lambda+mu)(Rg)(Rg)^T + mu|g|^2 I) (stiffnessDiagonal,
 * material.cc:290-304).  Neo-Hookean: the same blocks with the volumetric
 * part along the volume frame, V(mu(Rg)(Rg)^T + mu|g|^2 I +
 * lambda (Wg)(Wg)^T) -- the operator of op_element_range. */
static void element_one(orc_body *b, const double *at, size_t e, int diag) {
  double x[12], X[12], u[12], F[9], R[9];
  gather4(b, at, e, x);
  gather4(b, b->m.rest, e, X);
  for (int i = 0; i < 12; ++i) u[i] = x[i] - X[i];
  const double *dminv = &b->m.dminv[9 * e];
  const double V = b->m.volume[e];
  double energy = 0.0;
  if (b->snh[e]) {
    const double *ep = b->plastic_lane[e] ? &b->plastic[10 * e] : NULL;
    displacement_gradient(u, dminv, F);
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    if (ep) {
      double Fi[9], Fe[9];
      memcpy(Fi, ep, sizeof Fi);
      Fi[0] += 1.0;
      Fi[4] += 1.0;
      Fi[8] += 1.0;
      mul3(F, Fi, Fe);
      orc_element_rotation(Fe, R);
    } else {
      orc_element_rotation(F, R);
    }
    volume_frame(F, ep, &b->W[9 * e]);
    orc_snh_tet(u, dminv, V, b->lambda[e], b->mu[e], ep, &b->slot_f[12 * e], &energy);
  } else {
    j2_lin_t lin;
    corotated_eval(x, X, dminv, V, b->lambda[e], b->mu[e], &b->plastic[10 * e], b->yield[e],
                   b->hard[e], b->return_mapping, R, &b->slot_f[12 * e], &energy, NULL, NULL, &lin);
    if (b->plastic_lane[e]) {
      /* tetPointTrial (material.cc:1010-1022): loading past the margin;
       * the linearisation is the return map's when loading or elastic,
       * else the elastic law */
      const int loading = lin.plastic && lin.excess > b->margin;
      b->loading[e] = (uint8_t)loading;
      double *l = &b->lin[11 * e];
      memset(l, 0, 11 * sizeof(double));
      l[0] = 1.0;
      if (loading) {
        l[0] = lin.theta;
        l[1] = lin.theta_bar;
        memcpy(l + 2, lin.normal, sizeof lin.normal);
      }
    }
  }
  memcpy(&b->R[9 * e], R, sizeof R);
  b->tet_energy[e] = energy;
  if (b->tau[e] > 0.0) {
    /* fitDamping at the evaluation point (solver.cc:681-699); the damping
     * diagonal blocks of blockDamping / addDampingDiagonal (804-869) */
    rigid_fit(x, &b->fit[21 * e]);
    if (diag)
      for (int i = 0; i < 4; ++i)
        for (int axis = 0; axis < 3; ++axis) {
          double unit[12] = {0}, col[12];
          unit[3 * i + axis] = 1.0;
          damp_tet(b, e, unit, col);
          for (int r = 0; r < 3; ++r) b->slot_dd[36 * e + 9 * i + r * 3 + axis] = b->tau[e] * col[3 * i + r];
        }
  }
  if (diag) {
    double g[4][3];
    shape_gradients(dminv, g);
    const double *W = &b->W[9 * e];
    const double mu = b->mu[e], lam = b->lambda[e];
    if (!b->snh[e] && b->plastic_lane[e] && b->tangent) {
      /* stiffnessDiagonal of the consistent tangent blocks (material.cc:
       * 1033-1054, 290-304): R V((K + mu theta/3) g g^T + mu theta |g|^2 I
       * - 2 mu thetaBar (n g)(n g)^T) R^T, K = lambda + 2 mu / 3 */
      const double *l = &b->lin[11 * e];
      const double bulk = lam + 2.0 * mu / 3.0;
      for (int i = 0; i < 4; ++i) {
        double ng[3], B[9], RB[9];
        mv3(l + 2, g[i], ng);
        const double gg = g[i][0] * g[i][0] + g[i][1] * g[i][1] + g[i][2] * g[i][2];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c)
            B[r * 3 + c] = V * ((bulk + mu * l[0] / 3.0) * g[i][r] * g[i][c] +
                                (r == c ? mu * l[0] * gg : 0.0) - 2.0 * mu * l[1] * ng[r] * ng[c]);
        mul3(R, B, RB);
        mul3_nt(RB, R, &b->slot_d[36 * e + 9 * i]);
      }
      return;
    }
    double gs[4][3];
    plastic_gradients(b, e, g, gs);
    for (int i = 0; i < 4; ++i) {
      double rg[3], wg[3];
      mv3(R, gs[i], rg);
      mv3(W, g[i], wg);
      const double gg = gs[i][0] * gs[i][0] + gs[i][1] * gs[i][1] + gs[i][2] * gs[i][2];
      double *blk = &b->slot_d[36 * e + 9 * i];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          blk[r * 3 + c] = b->snh[e] ? V * (mu * rg[r] * rg[c] + lam * wg[r] * wg[c] + (r == c ? mu * gg : 0.0))
                                     : V * ((lam + mu) * rg[r] * rg[c] + (r == c ? mu * gg : 0.0));
    }
  }
}

static void element_range(void *p, size_t begin, size_t end) {
  elem_ctx *c = (elem_ctx *)p;
  for (size_t e = begin; e < end; ++e) element_one(c->b, c->at, e, c->diag);
}

typedef struct {
  orc_body *b;
  double scale;
  int diag;
} gather_ctx;

static void gather_range(void *p, size_t begin, size_t end) {
  gather_ctx *c = (gather_ctx *)p;
  orc_body *b = c->b;
  for (size_t n = begin; n < end; ++n) {
    double f[3] = {0, 0, 0}, d[9] = {0};
    for (uint32_t k = b->inc_start[n]; k < b->inc_start[n + 1]; ++k) {
      const uint32_t s = b->inc_slot[k];
      for (int r = 0; r < 3; ++r) f[r] += b->slot_f[3 * (size_t)s + r];
      if (c->diag)
        for (int i = 0; i < 9; ++i)
          d[i] += c->scale * b->slot_d[9 * (size_t)s + i] + (b->damped ? b->damp * b->slot_dd[9 * (size_t)s + i] : 0.0);
    }
    memcpy(&b->force[3 * n], f, sizeof f);
    if (c->diag) memcpy(&b->diag[9 * n], d, sizeof d);
  }
}

/* assembleElastic: returns the elastic energy */
static double assemble(orc_body *b, const double *at, double step,
                       uint32_t threads) {
  elem_ctx ec = {b, at, 1};
  parallel_for(b->m.tets, threads, element_range, &ec);
  gather_ctx gc = {b, step * step, 1};
  parallel_for(b->m.nodes, threads, gather_range, &gc);
  double energy = 0.0;
  for (size_t e = 0; e < b->m.tets; ++e) energy += b->tet_energy[e];
  return energy;
}

void orc_eval_forces(orc_body *b, const double *x, uint32_t threads,
                     double *tet_force, double *tet_energy, double *node_force,
                     double *rotations) {
  elem_ctx ec = {b, x, 0};
  parallel_for(b->m.tets, threads, element_range, &ec);
  gather_ctx gc = {b, 0.0, 0};
  parallel_for(b->m.nodes, threads, gather_range, &gc);
  if (tet_force) memcpy(tet_force, b->slot_f, 12 * (size_t)b->m.tets * sizeof(double));
  if (tet_energy) memcpy(tet_energy, b->tet_energy, (size_t)b->m.tets * sizeof(double));
  if (node_force) memcpy(node_force, b->force, 3 * (size_t)b->m.nodes * sizeof(double));
  if (rotations) memcpy(rotations, b->R, 9 * (size_t)b->m.tets * sizeof(double));
}

/* ---- operator ---- */

typedef struct {
  orc_body *b;
  const double *d;
  double *y;
  double inertia, scale;
} op_ctx;

/* Per tet: d'_j = R^T d_j, D' = sum d'_j g_j^T, sigma = lambda tr(D') I +
 * mu (D' + D'^T), y_i = R V sigma g_i -- equal to sum_j R K_ij R^T d_j with
 * K_ij = V(lambda g_i g_j^T + mu g_j g_i^T + mu (g_i.g_j) I)
 * (tetStiffness/block, material.cc:80-90, 209-234; applyStiffness 267-288).
 * Neo-Hookean tets keep the shear part and take the volumetric part along
 * the volume frame W: y_i = V(R mu (D' + D'^T) g_i + lambda (W:D) W g_i)
 * with D = sum d_j g_j^T.  At rest (R = W = I) both are the linear
 * stiffness; under stretch the second follows the stiffening of the
 * (lambda+mu)(J-1)^2 term along cof(F), which the rest stiffness misses. */
static void op_element_range(void *p, size_t begin, size_t end) {
  op_ctx *c = (op_ctx *)p;
  orc_body *b = c->b;
  for (size_t e = begin; e < end; ++e) {
    double d[12], g[4][3], gs[4][3], D[9] = {0}, Dw[9] = {0};
    gather4(b, c->d, e, d);
    shape_gradients(&b->m.dminv[9 * e], g);
    plastic_gradients(b, e, g, gs);
    const double *R = &b->R[9 * e];
    for (int j = 0; j < 4; ++j) {
      double dp[3];
      mtv3(R, &d[3 * j], dp);
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
          D[r * 3 + k] += dp[r] * gs[j][k];
          Dw[r * 3 + k] += d[3 * j + r] * g[j][k];
        }
    }
    const int snh = b->snh[e];
    const double tr = snh ? 0.0 : trace3(D);
    const double vol = snh ? b->lambda[e] * contract3(&b->W[9 * e], Dw) : 0.0;
    double s[9];
    if (!snh && b->plastic_lane[e] && b->tangent) {
      /* the consistent tangent (tangentOf, material.cc:636-647) on the
       * corotated strain sym(D') */
      double eps[9];
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) eps[r * 3 + k] = 0.5 * (D[r * 3 + k] + D[k * 3 + r]);
      j2_lin_t lin;
      const double *l = &b->lin[11 * e];
      lin.theta = l[0];
      lin.theta_bar = l[1];
      memcpy(lin.normal, l + 2, sizeof lin.normal);
      j2_tangent_stress(&lin, b->lambda[e], b->mu[e], eps, s);
    } else {
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
          s[r * 3 + k] = b->mu[e] * (D[r * 3 + k] + D[k * 3 + r]) +
                         (r == k ? b->lambda[e] * tr : 0.0);
    }
    const double V = b->m.volume[e];
    for (int i = 0; i < 4; ++i) {
      double t[3], y[3], w[3] = {0, 0, 0};
      mv3(s, gs[i], t);
      mv3(R, t, y);
      if (snh) mv3(&b->W[9 * e], g[i], w);
      for (int r = 0; r < 3; ++r) b->slot_f[12 * e + 3 * i + r] = V * (y[r] + vol * w[r]);
    }
    if (b->damped) {
      double out[12] = {0};
      if (b->tau[e] > 0.0) damp_tet(b, e, d, out);
      for (int k = 0; k < 12; ++k) b->slot_v[12 * e + k] = b->tau[e] * out[k];
    }
  }
}

static void op_gather_range(void *p, size_t begin, size_t end) {
  op_ctx *c = (op_ctx *)p;
  orc_body *b = c->b;
  for (size_t n = begin; n < end; ++n) {
    if (b->m.anchored[n]) {
      c->y[3 * n] = c->y[3 * n + 1] = c->y[3 * n + 2] = 0.0;
      continue;
    }
    double f[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    for (uint32_t k = b->inc_start[n]; k < b->inc_start[n + 1]; ++k) {
      const uint32_t s = b->inc_slot[k];
      for (int r = 0; r < 3; ++r) f[r] += b->slot_f[3 * (size_t)s + r];
      if (b->damped)
        for (int r = 0; r < 3; ++r) v[r] += b->slot_v[3 * (size_t)s + r];
    }
    const double m = c->inertia * b->m.mass[n];
    for (int r = 0; r < 3; ++r) c->y[3 * n + r] = m * c->d[3 * n + r] + c->scale * f[r] + b->damp * v[r];
  }
}

void orc_apply_system(orc_body *b, double inertia, double step,
                      const double *d, double *y, uint32_t threads) {
  op_ctx c = {b, d, y, inertia, step * step};
  parallel_for(b->m.tets, threads, op_element_range, &c);
  parallel_for(b->m.nodes, threads, op_gather_range, &c);
}

static double dot(orc_body *b, const double *x, const double *y,
                  uint32_t threads) {
  dot_ctx c = {x, y, NULL, b->m.nodes, b->partial};
  parallel_for(b->nblocks, threads > 1 && b->nblocks >= 64 ? threads : 1,
               dot_range, &c);
  double s = 0.0;
  for (size_t i = 0; i < b->nblocks; ++i) s += b->partial[i];
  return s;
}

static void precondition(orc_body *b) {
  for (size_t n = 0; n < b->m.nodes; ++n)
    mv3(&b->precond[9 * n], &b->residual[3 * n], &b->conditioned[3 * n]);
}

/* conjugateGradient (solver.cc:1904-2089) with the block-Jacobi
 * preconditioner; returns iterations, sets capped / relative residual. */
static uint32_t pcg(orc_body *b, double inertia, double step, uint32_t limit,
                    double tolerance, double *solution, uint32_t threads,
                    int *capped, double *relative) {
  const size_t N = b->m.nodes;
  const double reference = sqrt(dot(b, b->right, b->right, threads));
  orc_apply_system(b, inertia, step, solution, b->applied, threads);
  for (size_t n = 0; n < N; ++n)
    for (int r = 0; r < 3; ++r)
      b->residual[3 * n + r] =
          b->m.anchored[n] ? 0.0 : b->right[3 * n + r] - b->applied[3 * n + r];
  precondition(b);
  memcpy(b->direction, b->conditioned, 3 * N * sizeof(double));
  double squared = dot(b, b->residual, b->conditioned, threads);
  const double floor_ref = reference > DBL_MIN ? reference : DBL_MIN;
  const double target = tolerance * floor_ref;
  uint32_t taken = 0;
  for (uint32_t it = 0; it < limit; ++it) {
    const double length = dot(b, b->residual, b->residual, threads);
    if (sqrt(length) <= target) break;
    orc_apply_system(b, inertia, step, b->direction, b->applied, threads);
    const double curvature = dot(b, b->direction, b->applied, threads);
    if (!(curvature > 0.0) || !(squared > 0.0)) break;
    const double stride = squared / curvature;
    for (size_t i = 0; i < 3 * N; ++i) {
      solution[i] += stride * b->direction[i];
      b->residual[i] -= stride * b->applied[i];
    }
    precondition(b);
    const double next = dot(b, b->residual, b->conditioned, threads);
    const double blend = next / squared;
    squared = next;
    for (size_t i = 0; i < 3 * N; ++i)
      b->direction[i] = b->conditioned[i] + blend * b->direction[i];
    ++taken;
  }
  const double residual = sqrt(dot(b, b->residual, b->residual, threads));
  *capped = taken == limit && residual > target;
  *relative = residual / floor_ref;
  return taken;
}

/* integratorRule (solver.cc:531-545) and effectiveSystem (554-562) */
typedef struct {
  double am, af, beta, gamma;
} rule_t;

static rule_t rule_of(double spectral) {
  double s = spectral < 0.0 ? 0.0 : (spectral > 1.0 ? 1.0 : spectral);
  rule_t r;
  r.am = (2.0 * s - 1.0) / (s + 1.0);
  r.af = s / (s + 1.0);
  r.gamma = 0.5 - r.am + r.af;
  r.beta = 0.25 * (1.0 - r.am + r.af) * (1.0 - r.am + r.af);
  return r;
}

/* balance (solver.cc:3746-3806): right-hand side, inertia on the diagonal,
 * block-Jacobi inverse (guardPreconditioner, solver.cc:1574-1593). */
/* The damping force of balance (solver.cc:3746-3783): viscous = sum_e tau_e
 * D_e(v_w) at the alpha-point velocity v_w = v + lag (1 - gamma) a_n
 * [+ lag gamma a], lag = (1 - alpha_f) h, anchored nodes at rest; into
 * b->increment (free scratch between solves here). */
static void viscous_force(orc_body *b, const rule_t *rule, const double *taken, double *viscous) {
  const size_t N = b->m.nodes;
  double *vw = (double *)malloc(3 * N * sizeof(double) + 8);
  for (size_t n = 0; n < N; ++n)
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      vw[i] = b->m.anchored[n] ? 0.0
                               : b->v[i] + b->lag * (1.0 - rule->gamma) * b->carried[i] +
                                     (taken ? b->lag * rule->gamma * taken[i] : 0.0);
    }
  memset(viscous, 0, 3 * N * sizeof(double));
  for (size_t e = 0; e < b->m.tets; ++e) {
    if (!(b->tau[e] > 0.0)) continue;
    double ve[12], out[12];
    gather4(b, vw, e, ve);
    damp_tet(b, e, ve, out);
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 3; ++r) viscous[3 * (size_t)b->m.tet[4 * e + c] + r] += b->tau[e] * out[3 * c + r];
  }
  free(vw);
}

static void balance(orc_body *b, const rule_t *rule, double inertia,
                    const double *taken) {
  double *viscous = NULL;
  if (b->damped) {
    viscous = (double *)malloc(3 * (size_t)b->m.nodes * sizeof(double) + 8);
    viscous_force(b, rule, taken, viscous);
  }
  for (size_t n = 0; n < b->m.nodes; ++n) {
    double *P = &b->precond[9 * n];
    if (b->m.anchored[n]) {
      for (int r = 0; r < 3; ++r) {
        b->v[3 * n + r] = 0.0;
        b->a[3 * n + r] = 0.0;
        b->right[3 * n + r] = 0.0;
      }
      ident3(P);
      continue;
    }
    const double m = b->m.mass[n];
    for (int r = 0; r < 3; ++r) {
      double rhs = b->force[3 * n + r] + b->drive[3 * n + r] -
                   rule->am * m * b->carried[3 * n + r];
      if (taken) rhs -= inertia * m * taken[3 * n + r];
      if (viscous) rhs -= viscous[3 * n + r];
      b->right[3 * n + r] = rhs;
    }
    double *D = &b->diag[9 * n];
    D[0] += inertia * m;
    D[4] += inertia * m;
    D[8] += inertia * m;
    if (!(orc_det3(D) > 0.0) || !inv3(D, P)) ident3(P);
  }
  free(viscous);
}

/* predictorInverts (solver.cc:568-623) with FP64 determinants */
static int predictor_inverts(orc_body *b, const rule_t *rule, double inner) {
  const double carry = (0.5 - rule->beta) * inner * inner;
  for (size_t i = 0; i < 3 * (size_t)b->m.nodes; ++i) {
    const double shifted = b->x[i] + inner * b->v[i] + carry * b->a[i];
    b->weighted[i] = (1.0 - rule->af) * shifted + rule->af * b->x[i];
  }
  for (size_t e = 0; e < b->m.tets; ++e) {
    double w[12], X[12], u[12], F[9];
    gather4(b, b->weighted, e, w);
    gather4(b, b->m.rest, e, X);
    for (int i = 0; i < 12; ++i) u[i] = w[i] - X[i];
    displacement_gradient(u, &b->m.dminv[9 * e], F);
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    if (orc_det3(F) <= 0.0) return 1;
  }
  return 0;
}

/* commitPlastic (solver.cc:1260-1346) for corotated lanes and the Hencky
 * flow (plasticflow.cc:152-260) for neo-Hookean lanes, at final positions.
 * Returns the number of tets whose plastic state grew. */
static uint32_t commit_plastic(orc_body *b) {
  uint32_t grew = 0;
  for (size_t e = 0; e < b->m.tets; ++e) {
    if (!b->plastic_lane[e]) continue;
    double x[12], X[12], u[12], F[9];
    gather4(b, b->x, e, x);
    gather4(b, b->m.rest, e, X);
    for (int i = 0; i < 12; ++i) u[i] = x[i] - X[i];
    double *pl = &b->plastic[10 * e];
    if (b->snh[e]) {
      displacement_gradient(u, &b->m.dminv[9 * e], F);
      F[0] += 1.0;
      F[4] += 1.0;
      F[8] += 1.0;
      if (orc_hencky_flow(F, b->mu[e], b->yield[e], b->hard[e], b->pmax[e], pl,
                          &pl[9])) {
        b->yflag[e] = 1;
        ++grew;
      }
    } else {
      double R[9], f[12], next[10];
      orc_corotated_tet(x, X, &b->m.dminv[9 * e], b->m.volume[e], b->lambda[e],
                        b->mu[e], pl, b->yield[e], b->hard[e], R, f, NULL, next,
                        NULL);
      if (next[9] > pl[9]) {
        b->yflag[e] = 1;
        ++grew;
      }
      memcpy(pl, next, sizeof next);
    }
  }
  return grew;
}

/* One substep (solver.cc:3691-3950, 4351-4368, 5829-5849): predictor,
 * first solve, corrections ("anchored reads"), velocity and position
 * updates, plastic commit.  converge != 0 repeats the correction until the
 * increment is below 1e-12 of the acceleration (at most 64 passes): the root
 * of the substep residual, which the differentiable rollout needs. */
static double substep(orc_body *b, const orc_settings *s, const rule_t *rule, double inner,
                      orc_report *rep) {
  const size_t N = b->m.nodes;
  const uint32_t threads = s->threads;
  const double inertia = 1.0 - rule->am;
  const double step = sqrt((1.0 - rule->af) * rule->beta) * inner;
  const double carry = (0.5 - rule->beta) * inner * inner;
  const double reach = step * step;
  /* damping.scale and the alpha-point lag (solver.cc:3586, 3750) */
  b->damp = b->damped ? (1.0 - rule->af) * rule->gamma * inner : 0.0;
  b->lag = (1.0 - rule->af) * inner;
  for (size_t n = 0; n < N; ++n)
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      b->carried[i] = b->a[i];
      b->shifted[i] = b->x[i] + inner * b->v[i] + carry * b->a[i];
      b->weighted[i] = (1.0 - rule->af) * b->shifted[i] + rule->af * b->x[i];
      b->drive[i] = b->m.mass[n] * (double)s->gravity[r];
    }
  double elastic = assemble(b, b->weighted, step, threads);
  balance(b, rule, inertia, NULL);
  int capped = 0;
  double rel = 0.0;
  rep->cg_iterations += pcg(b, inertia, step, s->cg_iterations, s->tolerance, b->a, threads,
                            &capped, &rel);
  rep->cg_capped += (uint32_t)capped;
  if (rel > rep->cg_worst_relative) rep->cg_worst_relative = rel;
  ++rep->solves;
  const uint32_t reads = s->converge ? 64u : 3u;
  for (uint32_t read = 1; read < reads; ++read) {
    for (size_t i = 0; i < 3 * N; ++i) b->evaluated[i] = b->weighted[i] + reach * b->a[i];
    elastic = assemble(b, b->evaluated, step, threads);
    balance(b, rule, inertia, b->a);
    memset(b->increment, 0, 3 * N * sizeof(double));
    rep->cg_iterations += pcg(b, inertia, step, s->cg_iterations, s->tolerance, b->increment,
                              threads, &capped, &rel);
    rep->cg_capped += (uint32_t)capped;
    if (rel > rep->cg_worst_relative) rep->cg_worst_relative = rel;
    ++rep->solves;
    double change = 0.0, size = 0.0;
    for (size_t n = 0; n < N; ++n)
      if (!b->m.anchored[n])
        for (int r = 0; r < 3; ++r) {
          b->a[3 * n + r] += b->increment[3 * n + r];
          change = fmax(change, fabs(b->increment[3 * n + r]));
          size = fmax(size, fabs(b->a[3 * n + r]));
        }
    if (s->converge && change <= 1.0e-12 * size) break;
  }
  /* velocity (4351-4368) and position (5829-5849) */
  for (size_t n = 0; n < N; ++n) {
    if (b->m.anchored[n]) {
      for (int r = 0; r < 3; ++r) b->v[3 * n + r] = b->a[3 * n + r] = 0.0;
      continue;
    }
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      b->v[i] += inner * ((1.0 - rule->gamma) * b->carried[i] + rule->gamma * b->a[i]);
      b->x[i] = b->shifted[i] + rule->beta * inner * inner * b->a[i];
    }
  }
  const uint32_t grew = commit_plastic(b);
  rep->yield_events += grew;
  rep->plastic_growth |= grew ? 1u : 0u;
  return elastic;
}

/* ---- initialStress / applyInitialStress (material.cc:1060-1174) ---- */

static void cross_matrix(const double *v, double *m) {
  m[0] = 0.0; m[1] = -v[2]; m[2] = v[1];
  m[3] = v[2]; m[4] = 0.0; m[5] = -v[0];
  m[6] = -v[1]; m[7] = v[0]; m[8] = 0.0;
}

/* rotationDerivative for a tet (material.cc:1077-1118, spreadOf 690-709):
 * twist_i = R (sum_j K_ij [o_j]x) R^T + handed [f_i]x and
 * turn_i = spin [handed R g_i]x with spin = R (tr S I - S)^-1 R^T,
 * S = sym(R^T F); zero when F or tr S I - S is singular.  K_ij are the
 * elastic stiffness blocks (block, material.cc:80-90), o_j the corotated
 * offsets, f_i the element's corner forces.  twist, turn: 4 x 9 each. */
static int initial_stress_tet(const double *x, const double *X, const double *dminv, double V,
                              double lam, double mu, const double *R, const double *f,
                              double *twist, double *turn) {
  double g[4][3];
  shape_gradients(dminv, g);
  memset(twist, 0, 36 * sizeof(double));
  memset(turn, 0, 36 * sizeof(double));
  double F[9] = {0};
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) F[r * 3 + k] += x[3 * c + r] * g[c][k];
  if (fabs(orc_det3(F)) < 1.0e-12) return 0;
  double RtF[9], S[9], spread[9], inv[9];
  mul3_tn(R, F, RtF);
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) S[r * 3 + k] = 0.5 * (RtF[r * 3 + k] + RtF[k * 3 + r]);
  const double tr = trace3(S);
  for (int i = 0; i < 9; ++i) spread[i] = (i % 4 == 0 ? tr : 0.0) - S[i];
  if (!(orc_det3(spread) > 0.0) || !inv3(spread, inv)) return 0;
  double Ri[9], spin[9];
  mul3(R, inv, Ri);
  mul3_nt(Ri, R, spin);
  const double handed = orc_det3(R) < 0.0 ? -1.0 : 1.0;
  double xc[3] = {0, 0, 0}, Xc[3] = {0, 0, 0};
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 3; ++r) {
      xc[r] += 0.25 * x[3 * c + r];
      Xc[r] += 0.25 * X[3 * c + r];
    }
  double o[4][3];
  for (int c = 0; c < 4; ++c) {
    double arm[3];
    for (int r = 0; r < 3; ++r) arm[r] = x[3 * c + r] - xc[r];
    mtv3(R, arm, o[c]);
    for (int r = 0; r < 3; ++r) o[c][r] -= X[3 * c + r] - Xc[r];
  }
  for (int i = 0; i < 4; ++i) {
    double tw[9] = {0};
    for (int j = 0; j < 4; ++j) {
      double K[9], C[9], KC[9];
      const double gg = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
          K[r * 3 + k] = V * (lam * g[i][r] * g[j][k] + mu * g[j][r] * g[i][k] + (r == k ? mu * gg : 0.0));
      cross_matrix(o[j], C);
      mul3(K, C, KC);
      for (int k = 0; k < 9; ++k) tw[k] += KC[k];
    }
    double Rt[9], RtR[9], Cf[9];
    mul3(R, tw, Rt);
    mul3_nt(Rt, R, RtR);
    cross_matrix(&f[3 * i], Cf);
    for (int k = 0; k < 9; ++k) twist[9 * i + k] = RtR[k] + handed * Cf[k];
    double rg[3], Cg[9];
    mv3(R, g[i], rg);
    for (int r = 0; r < 3; ++r) rg[r] *= handed;
    cross_matrix(rg, Cg);
    mul3(spin, Cg, &turn[9 * i]);
  }
  return 1;
}

/* applyInitialStress (material.cc:1157-1174): turned = sum_k turn_k d_k,
 * twisted = sum_k twist_k^T d_k, y_i += (twist_i turned + turn_i^T twisted)/2 */
static void apply_initial_stress(const double *twist, const double *turn, const double *d, double *y) {
  double turned[3] = {0, 0, 0}, twisted[3] = {0, 0, 0};
  for (int k = 0; k < 4; ++k) {
    double a[3], b[3];
    mv3(&turn[9 * k], &d[3 * k], a);
    mtv3(&twist[9 * k], &d[3 * k], b);
    for (int r = 0; r < 3; ++r) {
      turned[r] += a[r];
      twisted[r] += b[r];
    }
  }
  for (int i = 0; i < 4; ++i) {
    double a[3], b[3];
    mv3(&twist[9 * i], turned, a);
    mtv3(&turn[9 * i], twisted, b);
    for (int r = 0; r < 3; ++r) y[3 * i + r] += 0.5 * (a[r] + b[r]);
  }
}

/* The system operator with the initial-stress term of the static Newton
 * (scratch.stressed, solver.cc:2623; applySystem 947-948): at positions x,
 * y = sum_e (R K R^T + initialStress) d over corotated tets (the elastic
 * stiffness, the element forces of the trial from the committed plastic
 * state), h = 1, no inertia; neo-Hookean tets add their operator only. */
void orc_stressed_apply(orc_body *b, const double *x, const double *d, double *y) {
  double *force = (double *)malloc(12 * (size_t)b->m.tets * sizeof(double) + 8);
  orc_eval_forces(b, x, 1, force, NULL, NULL, NULL);
  orc_apply_system(b, 0.0, 1.0, d, y, 1);
  for (size_t e = 0; e < b->m.tets; ++e) {
    if (b->snh[e]) continue;
    double xe[12], Xe[12], de[12], tw[36], tu[36], ye[12] = {0};
    gather4(b, x, e, xe);
    gather4(b, b->m.rest, e, Xe);
    gather4(b, d, e, de);
    if (!initial_stress_tet(xe, Xe, &b->m.dminv[9 * e], b->m.volume[e], b->lambda[e], b->mu[e],
                            &b->R[9 * e], &force[12 * e], tw, tu))
      continue;
    apply_initial_stress(tw, tu, de, ye);
    for (int c = 0; c < 4; ++c) {
      const uint32_t n = b->m.tet[4 * e + c];
      if (b->m.anchored[n]) continue;
      for (int r = 0; r < 3; ++r) y[3 * (size_t)n + r] += ye[3 * c + r];
    }
  }
  free(force);
}

/* The operator with stiffness damping at positions x (the fits and
 * rotations of an assembly at x): y = inertia M d + step^2 K d + damp
 * sum_e tau_e D_e d, anchored rows zero. */
void orc_damped_apply(orc_body *b, const double *x, double inertia, double step, double damp,
                      const double *d, double *y) {
  orc_eval_forces(b, x, 1, NULL, NULL, NULL, NULL);
  const double held = b->damp;
  b->damp = damp;
  orc_apply_system(b, inertia, step, d, y, 1);
  b->damp = held;
}

/* LOADING_MARGIN, PLASTIC_REDUCTION, PLASTIC_RESOLUTION (solver.cc:372-380) */
#define LOADING_MARGIN 1.0e-3
#define PLASTIC_REDUCTION 1.0e-2
#define PLASTIC_RESOLUTION 8.0
#define FLOAT_EPSILON 1.1920928955078125e-7

/* The J2 point law alone: stress and linearisation {theta, thetaBar,
 * normal[9]} of the return map at `strain` from the committed {eps_p,
 * alpha}; returns 1 when the point returns to the yield surface. */
int orc_j2_point(const double *strain, const double *plastic, double lambda, double mu,
                 double yield, double hardening, double *stress, double *lin11) {
  j2_lin_t lin;
  const int r = j2_return_map(strain, plastic, lambda, mu, yield, hardening, 1, stress, NULL, &lin);
  lin11[0] = lin.theta;
  lin11[1] = lin.theta_bar;
  memcpy(lin11 + 2, lin.normal, sizeof lin.normal);
  return r;
}

void orc_j2_tangent(const double *lin11, double lambda, double mu, const double *e, double *out) {
  j2_lin_t lin;
  lin.theta = lin11[0];
  lin.theta_bar = lin11[1];
  memcpy(lin.normal, lin11 + 2, sizeof lin.normal);
  j2_tangent_stress(&lin, lambda, mu, e, out);
}

static double assemble_plastic(orc_body *b, const double *at, double step, uint32_t threads,
                               double margin, int return_mapping, uint32_t *loading) {
  b->margin = margin;
  b->return_mapping = return_mapping;
  memset(b->loading, 0, b->m.tets);
  const double energy = assemble(b, at, step, threads);
  b->margin = 0.0;
  b->return_mapping = 1;
  uint32_t n = 0;
  for (size_t e = 0; e < b->m.tets; ++e) n += b->loading[e];
  *loading = n;
  return energy;
}

/* K_tangent d at positions x: assembled with the loading margin and the
 * consistent tangent bound (a Newton attempt's operator), y = K d (no
 * inertia, h = 1); returns the number of loading tets. */
uint32_t orc_tangent_apply(orc_body *b, const double *x, const double *d, double *y) {
  b->tangent = 1;
  uint32_t loading = 0;
  assemble_plastic(b, x, 1.0, 1, LOADING_MARGIN, 1, &loading);
  orc_apply_system(b, 0.0, 1.0, d, y, 1);
  b->tangent = 0;
  return loading;
}

/* assembleElastic with the loading margin and return-map switch; returns
 * the elastic energy and the number of loading tets (Assembly). */
static double right_squared(const orc_body *b) {
  double s = 0.0;
  for (size_t i = 0; i < 3 * (size_t)b->m.nodes; ++i) s += b->right[i] * b->right[i];
  return s;
}

/* One substep of a body with yielding corotated lanes: the damped plastic
 * Newton of deformable::step (solver.cc:3951-4250).  The unknown is the
 * acceleration `total`, evaluated at x_w + h_eff^2 total.  Attempt 0 solves
 * with the elastic stiffness from the previous acceleration; later attempts
 * solve K_tangent increment = residual with the consistent tangent and
 * halve the last increment while the residual norm grows (backtracking).
 * Converged when two successive evaluations load nothing or the position
 * change is below float resolution.  A predictor that loads nothing and a
 * residual that grows after the first solve take the elastic
 * three-solve scheme instead ("linear reads"), handing back to Newton if a
 * re-evaluation loads; a Newton that does not converge within `cap`
 * attempts falls back to the elastic scheme when that loads nothing and
 * reduces the residual 100-fold, else keeps the lowest-residual iterate. */
static double newton_substep(orc_body *b, const orc_settings *s, const rule_t *rule,
                             double inner, uint32_t cap, orc_report *rep) {
  const size_t N = b->m.nodes;
  const uint32_t threads = s->threads;
  const double inertia = 1.0 - rule->am;
  const double step = sqrt((1.0 - rule->af) * rule->beta) * inner;
  const double carry = (0.5 - rule->beta) * inner * inner;
  const double reach = step * step;
  const double travel = rule->beta * inner * inner;
  b->damp = b->damped ? (1.0 - rule->af) * rule->gamma * inner : 0.0;
  b->lag = (1.0 - rule->af) * inner;
  for (size_t n = 0; n < N; ++n)
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      b->carried[i] = b->a[i];
      b->shifted[i] = b->x[i] + inner * b->v[i] + carry * b->a[i];
      b->weighted[i] = (1.0 - rule->af) * b->shifted[i] + rule->af * b->x[i];
      b->drive[i] = b->m.mass[n] * (double)s->gravity[r];
    }
  double *total = b->total, *inc = b->increment;
  double elastic = 0.0;
  int capped = 0;
  double rel = 0.0;
#define EVALUATE()                                                           for (size_t i = 0; i < 3 * N; ++i) b->evaluated[i] = b->weighted[i] + reach * total[i]
#define SOLVE()                                                                            do {                                                                                       rep->cg_iterations += pcg(b, inertia, step, s->cg_iterations, s->tolerance, inc,                                    threads, &capped, &rel);                                       rep->cg_capped += (uint32_t)capped;                                                      if (rel > rep->cg_worst_relative) rep->cg_worst_relative = rel;                          ++rep->solves;                                                                         } while (0)
#define ADD_INCREMENT()                                                      for (size_t n = 0; n < N; ++n)                                               if (!b->m.anchored[n])                                                       for (int r = 0; r < 3; ++r) total[3 * n + r] += inc[3 * n + r]
  int linear_allowed = 1;
  for (;;) {
    memset(total, 0, 3 * N * sizeof(double));
    double accepted = INFINITY, opening = 0.0, lowest = INFINITY;
    uint32_t loading_before = 0, solves = 0, loading = 0;
    int converged = 0, predictor_unloaded = 0, handed_off = 0;
    for (uint32_t attempt = 0; attempt < cap; ++attempt) {
      if (attempt < 2) b->tangent = attempt != 0;
      EVALUATE();
      elastic = assemble_plastic(b, b->evaluated, step, threads, LOADING_MARGIN, 1, &loading);
      ++rep->plastic_evaluations;
      balance(b, rule, inertia, total);
      const double squared = right_squared(b);
      if (attempt != 0) {
        const double ranked = isfinite(squared) ? squared : INFINITY;
        if (attempt == 1 || ranked < lowest) {
          lowest = ranked;
          memcpy(b->lowest_it, total, 3 * N * sizeof(double));
        }
      }
      if (attempt == 0) predictor_unloaded = loading == 0;
      if (attempt == 1 && linear_allowed && predictor_unloaded && isfinite(squared) &&
          squared > accepted) {
        ++rep->plastic_linear_reads;
        b->tangent = 0;
        for (uint32_t read = 1; read < 3; ++read) {
          EVALUATE();
          uint32_t reread = 0;
          elastic = assemble_plastic(b, b->evaluated, step, threads, LOADING_MARGIN, 1, &reread);
          ++rep->plastic_evaluations;
          if (reread != 0) {
            handed_off = 1;
            break;
          }
          balance(b, rule, inertia, total);
          memset(inc, 0, 3 * N * sizeof(double));
          SOLVE();
          ++solves;
          ADD_INCREMENT();
        }
        converged = !handed_off;
        break;
      }
      if (attempt != 0 && !(squared <= accepted)) {
        for (size_t i = 0; i < 3 * N; ++i) {
          inc[i] *= 0.5;
          total[i] -= inc[i];
        }
        ++rep->plastic_backtracks;
        continue;
      }
      if (attempt != 0 && loading_before == 0 && loading == 0) {
        converged = 1;
        break;
      }
      if (attempt == 0) opening = squared;
      accepted = squared;
      loading_before = loading;
      if (attempt == 0)
        memcpy(inc, b->a, 3 * N * sizeof(double));
      else
        memset(inc, 0, 3 * N * sizeof(double));
      SOLVE();
      ++solves;
      double moved = 0.0, extent = 0.0;
      for (size_t n = 0; n < N; ++n) {
        double *d = &inc[3 * n];
        if (b->m.anchored[n]) {
          d[0] = d[1] = d[2] = 0.0;
          continue;
        }
        for (int r = 0; r < 3; ++r) total[3 * n + r] += d[r];
        const double *p = &b->evaluated[3 * n];
        moved = fmax(moved, sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
        extent = fmax(extent, sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
      }
      if (solves > 1 && travel * moved <= PLASTIC_RESOLUTION * FLOAT_EPSILON * extent) {
        converged = 1;
        break;
      }
    }
    rep->plastic_solves += solves;
    if (handed_off) {
      ++rep->plastic_handoffs;
      linear_allowed = 0;
      continue;
    }
    if (!converged) {
      b->tangent = 0;
      memset(total, 0, 3 * N * sizeof(double));
      int fell_back = 1;
      for (uint32_t read = 0; read < 3; ++read) {
        EVALUATE();
        uint32_t reread = 0;
        elastic = assemble_plastic(b, b->evaluated, step, threads, LOADING_MARGIN, read != 0, &reread);
        ++rep->plastic_evaluations;
        if (reread != 0) {
          fell_back = 0;
          break;
        }
        balance(b, rule, inertia, total);
        if (read == 0)
          memcpy(inc, b->a, 3 * N * sizeof(double));
        else
          memset(inc, 0, 3 * N * sizeof(double));
        SOLVE();
        ++rep->plastic_solves;
        ADD_INCREMENT();
      }
      if (fell_back) {
        EVALUATE();
        uint32_t reread = 0;
        elastic = assemble_plastic(b, b->evaluated, step, threads, LOADING_MARGIN, 1, &reread);
        ++rep->plastic_evaluations;
        balance(b, rule, inertia, total);
        const double squared = right_squared(b);
        fell_back = reread == 0 && squared <= PLASTIC_REDUCTION * PLASTIC_REDUCTION * opening;
      }
      if (fell_back) {
        ++rep->plastic_fallbacks;
        converged = 1;
      } else {
        ++rep->plastic_rejections;
        b->tangent = 1;
        memcpy(total, b->lowest_it, 3 * N * sizeof(double));
        EVALUATE();
        uint32_t reread = 0;
        elastic = assemble_plastic(b, b->evaluated, step, threads, LOADING_MARGIN, 1, &reread);
        ++rep->plastic_evaluations;
        balance(b, rule, inertia, total);
      }
    }
    rep->plastic_capped += converged ? 0u : 1u;
    memcpy(b->a, total, 3 * N * sizeof(double));
    break;
  }
#undef EVALUATE
#undef SOLVE
#undef ADD_INCREMENT
  b->tangent = 0;
  /* velocity (4351-4368) and position (5829-5849) */
  for (size_t n = 0; n < N; ++n) {
    if (b->m.anchored[n]) {
      for (int r = 0; r < 3; ++r) b->v[3 * n + r] = b->a[3 * n + r] = 0.0;
      continue;
    }
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      b->v[i] += inner * ((1.0 - rule->gamma) * b->carried[i] + rule->gamma * b->a[i]);
      b->x[i] = b->shifted[i] + rule->beta * inner * inner * b->a[i];
    }
  }
  const uint32_t grew = commit_plastic(b);
  rep->yield_events += grew;
  rep->plastic_growth |= grew ? 1u : 0u;
  return elastic;
}

/* One step of deformable::step for a non-floating tet body
 * (solver.cc:3472-5901 without contacts, damping, floating frame; the
 * elastic path 3895-3950 is used for every lane, see the header note). */
int orc_step(orc_body *b, const orc_settings *s, orc_report *rep) {
  memset(rep, 0, sizeof *rep);
  /* yield events of this step: tets whose plastic state grew in any substep */
  memset(b->yflag, 0, b->m.tets);
  const size_t N = b->m.nodes;
  const rule_t rule = rule_of(s->spectral_radius);
  const uint32_t configured = s->substeps > 1 ? s->substeps : 1;
  uint32_t substeps = configured;
  if (predictor_inverts(b, &rule, (double)s->dt / configured))
    substeps = configured > 4 ? configured : 4;
  const double inner = (double)s->dt / substeps;
  rep->substeps = substeps;
  double elastic = 0.0;
  /* newton = plasticPasses > 1 && anyYields (solver.cc:3517): here, bodies
   * with a yielding corotated lane (the return map runs in the trial) */
  const uint32_t passes = s->plastic_passes ? s->plastic_passes : 40u;
  const int newton = b->newton && passes > 1 && !s->converge;
  for (uint32_t pass = 0; pass < substeps; ++pass)
    elastic = newton ? newton_substep(b, s, &rule, inner, passes, rep) : substep(b, s, &rule, inner, rep);
  double kinetic = 0.0;
  for (size_t n = 0; n < N; ++n)
    kinetic += 0.5 * b->m.mass[n] *
               (b->v[3 * n] * b->v[3 * n] + b->v[3 * n + 1] * b->v[3 * n + 1] +
                b->v[3 * n + 2] * b->v[3 * n + 2]);
  rep->kinetic_energy = kinetic;
  rep->elastic_energy = elastic;
  for (size_t i = 0; i < 3 * N; ++i)
    if (!isfinite(b->x[i]) || !isfinite(b->v[i])) {
      rep->non_finite = 1;
      break;
    }
  return 0;
}

void orc_tet_stretch(orc_body *b, double *stretch) {
  for (size_t e = 0; e < b->m.tets; ++e) {
    double x[12], X[12], u[12], F[9];
    gather4(b, b->x, e, x);
    gather4(b, b->m.rest, e, X);
    for (int i = 0; i < 12; ++i) u[i] = x[i] - X[i];
    displacement_gradient(u, &b->m.dminv[9 * e], F);
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    stretch[e] = orc_stretch(F, NULL);
  }
}

/* overstretched (tearing.cc:469-483) on plain arrays: the stretch of every
 * tet of `tet` (4 node indices each, positive orientation at rest) at the
 * positions `x`, with the rest inverse rebuilt from `rest` as the lattice
 * build makes it */
void orc_stretch_tets(const uint32_t *tet, size_t tets, const double *rest, const double *x,
                      double *stretch) {
  for (size_t e = 0; e < tets; ++e) {
    const uint32_t *tn = &tet[4 * e];
    double shape[9], dminv[9], u[12], F[9];
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r)
        shape[r * 3 + c] = rest[3 * tn[c + 1] + r] - rest[3 * tn[0] + r];
    inv3(shape, dminv);
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) u[3 * k + r] = x[3 * tn[k] + r] - rest[3 * tn[k] + r];
    displacement_gradient(u, dminv, F);
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    stretch[e] = orc_stretch(F, NULL);
  }
}

/* ------------------------------------------------------------------ */
/* Tearing topology (tearing.cc:180-513 restated for tets)              */
/* ------------------------------------------------------------------ */

typedef struct {
  uint32_t nodes, tets, capacity, copies;
  uint32_t *tet;
  double *rest, *x;
  uint8_t *anchored;
  uint32_t *origin;
  int overflow;
} topo_t;

/* tets containing node, ascending; returns the count (out has room for tets) */
static uint32_t topo_fan(const topo_t *t, uint32_t node, uint32_t *out) {
  uint32_t n = 0;
  for (uint32_t e = 0; e < t->tets; ++e) {
    const uint32_t *c = &t->tet[4 * (size_t)e];
    if (c[0] == node || c[1] == node || c[2] == node || c[3] == node) out[n++] = e;
  }
  return n;
}

static uint32_t topo_shared(const topo_t *t, uint32_t a, uint32_t b) {
  uint32_t n = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) n += t->tet[4 * (size_t)a + i] == t->tet[4 * (size_t)b + j];
  return n;
}

static uint32_t uf_root(uint32_t *parent, uint32_t x) {
  while (parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = parent[x];
  }
  return x;
}

/* groupsOf (tearing.cc:92-129) with face sharing = 3 common nodes */
static void topo_groups(const topo_t *t, const uint32_t *fan, uint32_t n, const uint32_t *sides,
                        uint32_t *group) {
  uint32_t *parent = (uint32_t *)malloc((n + 1) * sizeof(uint32_t));
  uint32_t *id = (uint32_t *)malloc((n + 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) {
    parent[i] = i;
    id[i] = UINT32_MAX;
  }
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j)
      if (sides[i] == sides[j] && topo_shared(t, fan[i], fan[j]) >= 3u) {
        const uint32_t a = uf_root(parent, i), b = uf_root(parent, j);
        if (a != b) parent[a > b ? a : b] = a < b ? a : b;
      }
  uint32_t next = 0;
  for (uint32_t side = 0; side < 2; ++side)
    for (uint32_t i = 0; i < n; ++i) {
      if (sides[i] != side) continue;
      const uint32_t r = uf_root(parent, i);
      if (id[r] == UINT32_MAX) id[r] = next++;
      group[i] = id[r];
    }
  free(parent);
  free(id);
}

/* reachable(seeds, pivot, MIN_PIECE_CELLS = 2) >= 2 */
static int topo_whole(const topo_t *t, const uint32_t *seeds, uint32_t nseeds, uint32_t pivot) {
  if (nseeds >= 2) return 1;
  const uint32_t e = seeds[0];
  const uint32_t *c = &t->tet[4 * (size_t)e];
  for (int f = 0; f < 4; ++f) {
    uint32_t fc[3];
    int w = 0;
    for (int k = 0; k < 4; ++k)
      if (k != f) fc[w++] = c[k];
    if (fc[0] == pivot || fc[1] == pivot || fc[2] == pivot) continue;
    for (uint32_t o = 0; o < t->tets; ++o) {
      if (o == e) continue;
      const uint32_t *d = &t->tet[4 * (size_t)o];
      int hits = 0;
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 4; ++k) hits += d[k] == fc[a];
      if (hits == 3) return 1;
    }
  }
  return 0;
}

static int topo_split(topo_t *t, uint32_t node, const uint32_t *fan, uint32_t n,
                      const uint32_t *group) {
  uint32_t count = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (group[i] + 1 > count) count = group[i] + 1;
  if (count < 2 || t->anchored[node]) return 0;
  for (uint32_t piece = 1; piece < count; ++piece) {
    if (t->nodes >= t->capacity) {
      t->overflow = 1;
      return 0;
    }
    const uint32_t copy = t->nodes++;
    for (int k = 0; k < 3; ++k) {
      t->rest[3 * (size_t)copy + k] = t->rest[3 * (size_t)node + k];
      t->x[3 * (size_t)copy + k] = t->x[3 * (size_t)node + k];
    }
    t->anchored[copy] = 0;
    t->origin[t->copies++] = node;
    for (uint32_t i = 0; i < n; ++i) {
      if (group[i] != piece) continue;
      uint32_t *c = &t->tet[4 * (size_t)fan[i]];
      for (int k = 0; k < 4; ++k)
        if (c[k] == node) c[k] = copy;
    }
  }
  return 1;
}

static void canonical_dir(double *d) {
  int best = 0;
  for (int k = 1; k < 3; ++k)
    if (fabs(d[k]) > fabs(d[best])) best = k;
  if (d[best] < 0.0)
    for (int k = 0; k < 3; ++k) d[k] = -d[k];
}

uint32_t orc_tear_topology(uint32_t nodes, uint32_t tets, uint32_t *tet, double *rest, double *x,
                           uint8_t *anchored, uint32_t capacity, const uint32_t *torn,
                           uint32_t ntorn, uint32_t *origin) {
  topo_t t = {nodes, tets, capacity, 0, tet, rest, x, anchored, origin, 0};
  uint32_t *fan = (uint32_t *)malloc((tets + 1) * sizeof(uint32_t));
  uint32_t *sides = (uint32_t *)malloc((tets + 1) * sizeof(uint32_t));
  uint32_t *group = (uint32_t *)malloc((tets + 1) * sizeof(uint32_t));
  uint32_t *seeds = (uint32_t *)malloc((tets + 1) * sizeof(uint32_t));
  for (uint32_t q = 0; q < ntorn; ++q) {
    const uint32_t cell = torn[q];
    if (cell >= tets) continue;
    uint32_t corners[4];
    memcpy(corners, &tet[4 * (size_t)cell], sizeof corners);
    /* crack normal: top eigenvector of F^T F, F = ds Ds^-1 */
    double Ds[9], ds[9], Dinv[9], F[9], across[3];
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) {
        Ds[r * 3 + k] = rest[3 * (size_t)corners[k + 1] + r] - rest[3 * (size_t)corners[0] + r];
        ds[r * 3 + k] = x[3 * (size_t)corners[k + 1] + r] - x[3 * (size_t)corners[0] + r];
      }
    inv3(Ds, Dinv);
    mul3(ds, Dinv, F);
    orc_stretch(F, across);
    canonical_dir(across);
    double centre[3] = {0, 0, 0};
    for (int c = 0; c < 4; ++c)
      for (int k = 0; k < 3; ++k) centre[k] += rest[3 * (size_t)corners[c] + k] / 4.0;
    double dist[4];
    int order[4] = {0, 1, 2, 3};
    for (int c = 0; c < 4; ++c) {
      double d = 0.0;
      for (int k = 0; k < 3; ++k) d += (rest[3 * (size_t)corners[c] + k] - centre[k]) * across[k];
      dist[c] = fabs(d);
    }
    for (int i = 0; i < 4; ++i)
      for (int j = i + 1; j < 4; ++j) {
        const int a = order[i], b = order[j];
        if (dist[b] < dist[a] || (dist[b] == dist[a] && corners[b] < corners[a])) {
          order[i] = b;
          order[j] = a;
        }
      }
    for (int oi = 0; oi < 4; ++oi) {
      const uint32_t node = corners[order[oi]];
      if (anchored[node]) continue;
      const uint32_t n = topo_fan(&t, node, fan);
      int low = 0, high = 0;
      for (uint32_t i = 0; i < n; ++i) {
        double offset = 0.0;
        const uint32_t *c = &tet[4 * (size_t)fan[i]];
        for (int a = 0; a < 4; ++a)
          for (int k = 0; k < 3; ++k) offset += (rest[3 * (size_t)c[a] + k] - rest[3 * (size_t)node + k]) * across[k];
        sides[i] = offset > 0.0 ? 1u : 0u;
        low |= sides[i] == 0u;
        high |= sides[i] == 1u;
      }
      if (!low || !high) continue;
      topo_groups(&t, fan, n, sides, group);
      uint32_t count = 0;
      for (uint32_t i = 0; i < n; ++i)
        if (group[i] + 1 > count) count = group[i] + 1;
      int whole = 1;
      for (uint32_t piece = 0; piece < count && whole; ++piece) {
        uint32_t ns = 0;
        for (uint32_t i = 0; i < n; ++i)
          if (group[i] == piece) seeds[ns++] = fan[i];
        whole = topo_whole(&t, seeds, ns, node);
      }
      if (whole && topo_split(&t, node, fan, n, group)) break;
    }
  }
  if (t.copies) {
    /* splitPinched: nodes of tets touching a split node, ascending */
    uint8_t *near = (uint8_t *)calloc(t.nodes + 1, 1);
    uint8_t *cand = (uint8_t *)calloc(t.nodes + 1, 1);
    for (uint32_t i = 0; i < t.copies; ++i) {
      near[nodes + i] = 1;
      near[origin[i]] = 1;
    }
    for (uint32_t e = 0; e < tets; ++e) {
      const uint32_t *c = &tet[4 * (size_t)e];
      if (near[c[0]] || near[c[1]] || near[c[2]] || near[c[3]])
        for (int k = 0; k < 4; ++k) cand[c[k]] = 1;
    }
    const uint32_t limit = t.nodes;
    for (uint32_t node = 0; node < limit; ++node) {
      if (!cand[node]) continue;
      const uint32_t n = topo_fan(&t, node, fan);
      if (n < 2) continue;
      for (uint32_t i = 0; i < n; ++i) sides[i] = 0;
      topo_groups(&t, fan, n, sides, group);
      (void)topo_split(&t, node, fan, n, group);
    }
    free(near);
    free(cand);
  }
  free(fan);
  free(sides);
  free(group);
  free(seeds);
  return t.overflow ? 0u : t.nodes;
}

uint32_t orc_body_nodes(const orc_body *b) { return b->m.nodes; }

void orc_body_topology(const orc_body *b, uint32_t *tet, double *rest, double *mass,
                       uint8_t *anchored) {
  if (tet) memcpy(tet, b->m.tet, 4 * (size_t)b->m.tets * sizeof(uint32_t));
  if (rest) memcpy(rest, b->m.rest, 3 * (size_t)b->m.nodes * sizeof(double));
  if (mass) memcpy(mass, b->m.mass, (size_t)b->m.nodes * sizeof(double));
  if (anchored) memcpy(anchored, b->m.anchored, b->m.nodes);
}

static void *grow(void *p, size_t bytes) {
  void *q = realloc(p, bytes ? bytes : 1);
  return q;
}

uint32_t orc_body_tear(orc_body *b, const uint32_t *torn, uint32_t ntorn) {
  const uint32_t N = b->m.nodes, T = b->m.tets;
  const uint32_t cap = N + 64u * (ntorn + 1u);
  double *rest = (double *)malloc(3 * (size_t)cap * sizeof(double));
  double *x = (double *)malloc(3 * (size_t)cap * sizeof(double));
  uint8_t *anchored = (uint8_t *)malloc(cap);
  uint32_t *origin = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
  memcpy(rest, b->m.rest, 3 * (size_t)N * sizeof(double));
  memcpy(x, b->x, 3 * (size_t)N * sizeof(double));
  memcpy(anchored, b->m.anchored, N);
  const uint32_t n2 = orc_tear_topology(N, T, b->m.tet, rest, x, anchored, cap, torn, ntorn, origin);
  const uint32_t added = n2 > N ? n2 - N : 0;
  if (added) {
    free(b->m.rest);
    b->m.rest = rest;
    free(b->m.anchored);
    b->m.anchored = anchored;
    free(b->x);
    b->x = x;
    b->v = (double *)grow(b->v, 3 * (size_t)n2 * sizeof(double));
    b->a = (double *)grow(b->a, 3 * (size_t)n2 * sizeof(double));
    for (uint32_t i = 0; i < added; ++i)
      for (int k = 0; k < 3; ++k) {
        b->v[3 * (size_t)(N + i) + k] = b->v[3 * (size_t)origin[i] + k];
        b->a[3 * (size_t)(N + i) + k] = b->a[3 * (size_t)origin[i] + k];
      }
    b->m.nodes = n2;
    /* recomputeMasses: rho V / 4 per corner */
    b->m.mass = (double *)grow(b->m.mass, (size_t)n2 * sizeof(double));
    memset(b->m.mass, 0, (size_t)n2 * sizeof(double));
    for (size_t e = 0; e < T; ++e)
      for (int c = 0; c < 4; ++c)
        b->m.mass[b->m.tet[4 * e + c]] += b->rho[e] * b->m.volume[e] * 0.25;
    /* incidence and node scratch for the new node count */
    free(b->inc_start);
    free(b->inc_slot);
    b->inc_start = (uint32_t *)calloc((size_t)n2 + 1, sizeof(uint32_t));
    for (size_t q = 0; q < 4 * (size_t)T; ++q) ++b->inc_start[b->m.tet[q] + 1];
    for (size_t n = 0; n < n2; ++n) b->inc_start[n + 1] += b->inc_start[n];
    b->inc_slot = (uint32_t *)malloc(4 * (size_t)T * sizeof(uint32_t) + 4);
    uint32_t *cursor = (uint32_t *)malloc((size_t)n2 * sizeof(uint32_t) + 4);
    memcpy(cursor, b->inc_start, (size_t)n2 * sizeof(uint32_t));
    for (size_t q = 0; q < 4 * (size_t)T; ++q) b->inc_slot[cursor[b->m.tet[q]]++] = (uint32_t)q;
    free(cursor);
    double **vecs[] = {&b->force, &b->right, &b->weighted, &b->shifted, &b->carried, &b->drive,
                       &b->evaluated, &b->residual, &b->direction, &b->applied,
                       &b->conditioned, &b->increment, &b->unconstrained};
    for (size_t i = 0; i < sizeof vecs / sizeof vecs[0]; ++i) {
      free(*vecs[i]);
      *vecs[i] = (double *)calloc(3 * (size_t)n2, sizeof(double));
    }
    free(b->diag);
    free(b->precond);
    b->diag = (double *)calloc(9 * (size_t)n2, sizeof(double));
    b->precond = (double *)calloc(9 * (size_t)n2, sizeof(double));
    free(b->partial);
    b->nblocks = ((size_t)n2 + DOT_BLOCK - 1) / DOT_BLOCK;
    b->partial = (double *)calloc(b->nblocks + 1, sizeof(double));
  } else {
    free(rest);
    free(x);
    free(anchored);
  }
  free(origin);
  return added;
}

/* ------------------------------------------------------------------ */
/* Per-element Lame parameters and rest positions in FP64               */
/* ------------------------------------------------------------------ */

void orc_body_set_lame(orc_body *b, const double *lambda, const double *mu) {
  memcpy(b->lambda, lambda, b->m.tets * sizeof(double));
  memcpy(b->mu, mu, b->m.tets * sizeof(double));
}

void orc_body_get_lame(const orc_body *b, double *lambda, double *mu) {
  memcpy(lambda, b->lambda, b->m.tets * sizeof(double));
  memcpy(mu, b->mu, b->m.tets * sizeof(double));
}

/* New rest positions: rest inverse, volume det(Dm)/6 and lumped masses
 * rho V / 4 recomputed in FP64 (buildMesh, solver.cc:2352-2364). */
int orc_body_set_rest(orc_body *b, const double *X) {
  const size_t N = b->m.nodes, T = b->m.tets;
  memcpy(b->m.rest, X, 3 * N * sizeof(double));
  memset(b->m.mass, 0, N * sizeof(double));
  for (size_t e = 0; e < T; ++e) {
    const uint32_t *t = &b->m.tet[4 * e];
    double Ds[9];
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) Ds[r * 3 + k] = X[3 * t[k + 1] + r] - X[3 * t[0] + r];
    const double det = orc_det3(Ds);
    if (!(det > 0.0) || !inv3(Ds, &b->m.dminv[9 * e])) return -1;
    b->m.volume[e] = det / 6.0;
    for (int c = 0; c < 4; ++c) b->m.mass[t[c]] += b->rho[e] * b->m.volume[e] * 0.25;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Adjoint of the converged rollout (no counterpart in the reference:   */
/* the differentiation pattern of this build), neo-Hookean, elastic     */
/* ------------------------------------------------------------------ */

/* P(F) = mu (F - cof F) + (lambda + mu)(J - 1) cof F, the SNH first Piola
 * stress in F form (equal to the H form of orc_snh_tet). */
static void snh_P(const double *F, double lam, double mu, double *P) {
  double C[9];
  cof3(F, C);
  const double J = orc_det3(F);
  for (int i = 0; i < 9; ++i) P[i] = mu * (F[i] - C[i]) + (lam + mu) * (J - 1.0) * C[i];
}

/* d cof(F)[D]: the columns of cof F are f1 x f2, f2 x f0, f0 x f1 (0-based
 * columns f of F), so each column's derivative is two cross products. */
static void cof_dir(const double *F, const double *D, double *dC) {
  for (int c = 0; c < 3; ++c) {
    const int a = (c + 1) % 3, q = (c + 2) % 3;
    const double fa[3] = {F[a], F[3 + a], F[6 + a]}, fb[3] = {F[q], F[3 + q], F[6 + q]};
    const double da[3] = {D[a], D[3 + a], D[6 + a]}, db[3] = {D[q], D[3 + q], D[6 + q]};
    const double x1[3] = {da[1] * fb[2] - da[2] * fb[1], da[2] * fb[0] - da[0] * fb[2],
                          da[0] * fb[1] - da[1] * fb[0]};
    const double x2[3] = {fa[1] * db[2] - fa[2] * db[1], fa[2] * db[0] - fa[0] * db[2],
                          fa[0] * db[1] - fa[1] * db[0]};
    for (int r = 0; r < 3; ++r) dC[r * 3 + c] = x1[r] + x2[r];
  }
}

/* dP = dP/dF [D] = mu (D - dcof) + (lambda + mu)((cof F : D) cof F + (J - 1) dcof) */
static void snh_dP(const double *F, const double *D, double lam, double mu, double *dP) {
  double C[9], dC[9];
  cof3(F, C);
  cof_dir(F, D, dC);
  const double J = orc_det3(F), cd = contract3(C, D);
  for (int i = 0; i < 9; ++i)
    dP[i] = mu * (D[i] - dC[i]) + (lam + mu) * (cd * C[i] + (J - 1.0) * dC[i]);
}

static void tet_F(const orc_body *b, const double *pos, size_t e, double *F) {
  double x[12], X[12], u[12];
  gather4(b, pos, e, x);
  gather4(b, b->m.rest, e, X);
  for (int i = 0; i < 12; ++i) u[i] = x[i] - X[i];
  displacement_gradient(u, &b->m.dminv[9 * e], F);
  F[0] += 1.0;
  F[4] += 1.0;
  F[8] += 1.0;
}

/* y = K(pos) d, K the SNH Hessian of the elastic energy (all rows). */
void orc_hessian_apply(orc_body *b, const double *pos, const double *d, double *y) {
  const size_t N = b->m.nodes, T = b->m.tets;
  memset(y, 0, 3 * N * sizeof(double));
  for (size_t e = 0; e < T; ++e) {
    double F[9], de[12], D[9], dP[9], g[4][3];
    tet_F(b, pos, e, F);
    gather4(b, d, e, de);
    displacement_gradient(de, &b->m.dminv[9 * e], D);
    snh_dP(F, D, b->lambda[e], b->mu[e], dP);
    shape_gradients(&b->m.dminv[9 * e], g);
    const double V = b->m.volume[e];
    for (int i = 0; i < 4; ++i) {
      double t[3];
      mv3(dP, g[i], t);
      const uint32_t n = b->m.tet[4 * e + i];
      for (int r = 0; r < 3; ++r) y[3 * n + r] += V * t[r];
    }
  }
}

/* Solves (inertia M + reach K(pos)) z = rhs on the free rows (anchored rows
 * and entries zero) by block-Jacobi PCG to `tolerance` relative. */
static uint32_t adjoint_solve(orc_body *b, const double *pos, double inertia, double reach,
                              const double *rhs, double *z, double tolerance) {
  const size_t N = b->m.nodes, T = b->m.tets;
  double *blk = (double *)calloc(9 * N, sizeof(double));
  for (size_t e = 0; e < T; ++e) {
    double F[9], g[4][3];
    tet_F(b, pos, e, F);
    shape_gradients(&b->m.dminv[9 * e], g);
    const double V = b->m.volume[e];
    for (int i = 0; i < 4; ++i) {
      const uint32_t n = b->m.tet[4 * e + i];
      for (int c = 0; c < 3; ++c) {
        double D[9] = {0}, dP[9], t[3];
        for (int k = 0; k < 3; ++k) D[c * 3 + k] = g[i][k];
        snh_dP(F, D, b->lambda[e], b->mu[e], dP);
        mv3(dP, g[i], t);
        for (int r = 0; r < 3; ++r) blk[9 * n + r * 3 + c] += reach * V * t[r];
      }
    }
  }
  double *inv = (double *)calloc(9 * N, sizeof(double));
  for (size_t n = 0; n < N; ++n) {
    double *B = &blk[9 * n];
    B[0] += inertia * b->m.mass[n];
    B[4] += inertia * b->m.mass[n];
    B[8] += inertia * b->m.mass[n];
    if (b->m.anchored[n] || !(orc_det3(B) > 0.0) || !inv3(B, &inv[9 * n])) ident3(&inv[9 * n]);
  }
  double *r = (double *)calloc(3 * N, sizeof(double)), *p = (double *)calloc(3 * N, sizeof(double));
  double *q = (double *)calloc(3 * N, sizeof(double)), *w = (double *)calloc(3 * N, sizeof(double));
  double *Kp = (double *)calloc(3 * N, sizeof(double));
  memset(z, 0, 3 * N * sizeof(double));
  double ref = 0.0;
  for (size_t n = 0; n < N; ++n)
    for (int c = 0; c < 3; ++c) {
      r[3 * n + c] = b->m.anchored[n] ? 0.0 : rhs[3 * n + c];
      ref += r[3 * n + c] * r[3 * n + c];
    }
  ref = sqrt(ref);
  for (size_t n = 0; n < N; ++n) mv3(&inv[9 * n], &r[3 * n], &w[3 * n]);
  memcpy(p, w, 3 * N * sizeof(double));
  double rz = 0.0;
  for (size_t i = 0; i < 3 * N; ++i) rz += r[i] * w[i];
  uint32_t it = 0;
  for (; it < 100000u && ref > 0.0; ++it) {
    double rr = 0.0;
    for (size_t i = 0; i < 3 * N; ++i) rr += r[i] * r[i];
    if (sqrt(rr) <= tolerance * ref) break;
    orc_hessian_apply(b, pos, p, Kp);
    double pq = 0.0;
    for (size_t n = 0; n < N; ++n)
      for (int c = 0; c < 3; ++c) {
        const size_t i = 3 * n + c;
        q[i] = b->m.anchored[n] ? 0.0 : inertia * b->m.mass[n] * p[i] + reach * Kp[i];
        pq += p[i] * q[i];
      }
    if (!(pq > 0.0)) break;
    const double alpha = rz / pq;
    for (size_t i = 0; i < 3 * N; ++i) {
      z[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    for (size_t n = 0; n < N; ++n) mv3(&inv[9 * n], &r[3 * n], &w[3 * n]);
    double next = 0.0;
    for (size_t i = 0; i < 3 * N; ++i) next += r[i] * w[i];
    const double beta = next / rz;
    rz = next;
    for (size_t i = 0; i < 3 * N; ++i) p[i] = w[i] + beta * p[i];
  }
  free(blk);
  free(inv);
  free(r);
  free(p);
  free(q);
  free(w);
  free(Kp);
  return it;
}

int orc_rollout(orc_body *b, const orc_settings *s, uint32_t steps, uint32_t every) {
  const size_t N = b->m.nodes;
  if (every == 0) every = 1;
  free(b->ckpt);
  free(b->final_state);
  const uint32_t count = (steps + every - 1) / every + 1;
  b->ckpt = (double *)malloc((size_t)count * 9 * N * sizeof(double));
  b->final_state = (double *)malloc(9 * N * sizeof(double));
  b->ckpt_steps = steps;
  b->ckpt_every = every;
  b->ckpt_settings = *s;
  b->ckpt_settings.converge = 1;
  for (uint32_t t = 0; t < steps; ++t) {
    if (t % every == 0) {
      double *c = &b->ckpt[(size_t)(t / every) * 9 * N];
      memcpy(c, b->x, 3 * N * sizeof(double));
      memcpy(c + 3 * N, b->v, 3 * N * sizeof(double));
      memcpy(c + 6 * N, b->a, 3 * N * sizeof(double));
    }
    orc_report rep;
    orc_step(b, &b->ckpt_settings, &rep);
    if (rep.non_finite) return -1;
  }
  memcpy(b->final_state, b->x, 3 * N * sizeof(double));
  memcpy(b->final_state + 3 * N, b->v, 3 * N * sizeof(double));
  memcpy(b->final_state + 6 * N, b->a, 3 * N * sizeof(double));
  return 0;
}

/* Reverse pass of one converged substep.  Forward (free node i):
 *   x_w = x_n + (1-af)(h v_n + c a_n) + reach a',  c = (1/2-beta) h^2,
 *   r(a') = inertia M a' + am M a_n - f(x_w) - M g = 0,
 *   x' = x_n + h v_n + c a_n + beta h^2 a',  v' = v_n + h((1-gamma) a_n + gamma a');
 * anchored nodes keep x and have v = a = 0.  With lambda = J^-1 abar_total,
 * J = inertia M + reach K(x_w) (implicit function theorem):
 *   xbar_n = xbar' - K lambda,  vbar_n = h xbar' + vbar' - (1-af) h K lambda,
 *   abar_n = c xbar' + h(1-gamma) vbar' - am M lambda - (1-af) c K lambda,
 * and the parameters get -d/dtheta [sum_e V_e P_e : D_e(lambda)] minus the
 * mass terms (inertia a' + am a_n - g) . lambda dm/dtheta. */
static void adjoint_substep(orc_body *b, const rule_t *rule, double inner, const float *gravity,
                            const double *xn, const double *vn, const double *an,
                            const double *anext, double *xb, double *vb, double *ab,
                            double *lam_bar, double *mu_bar, double *X_bar, double tolerance) {
  const size_t N = b->m.nodes, T = b->m.tets;
  const double inertia = 1.0 - rule->am;
  const double carry = (0.5 - rule->beta) * inner * inner;
  const double reach = (1.0 - rule->af) * rule->beta * inner * inner;
  double *xw = (double *)malloc(3 * N * sizeof(double));
  double *seed = (double *)calloc(3 * N, sizeof(double));
  double *lam = (double *)calloc(3 * N, sizeof(double));
  double *Kl = (double *)calloc(3 * N, sizeof(double));
  double *w = (double *)calloc(N, sizeof(double));
  for (size_t i = 0; i < 3 * N; ++i)
    xw[i] = xn[i] + (1.0 - rule->af) * (inner * vn[i] + carry * an[i]) + reach * anext[i];
  for (size_t n = 0; n < N; ++n)
    if (!b->m.anchored[n])
      for (int r = 0; r < 3; ++r) {
        const size_t i = 3 * n + r;
        seed[i] = ab[i] + rule->beta * inner * inner * xb[i] + rule->gamma * inner * vb[i];
      }
  adjoint_solve(b, xw, inertia, reach, seed, lam, tolerance);
  orc_hessian_apply(b, xw, lam, Kl);
  for (size_t n = 0; n < N; ++n) {
    if (b->m.anchored[n]) continue;
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      w[n] += (inertia * anext[i] + rule->am * an[i] - (double)gravity[r]) * lam[i];
    }
  }
  /* parameters, at x_w with D = sum_j lambda_j g_j^T */
  for (size_t e = 0; e < T; ++e) {
    double F[9], C[9], le[12], D[9], P[9], dP[9];
    tet_F(b, xw, e, F);
    cof3(F, C);
    const double J = orc_det3(F);
    gather4(b, lam, e, le);
    const double *dm = &b->m.dminv[9 * e];
    displacement_gradient(le, dm, D);
    const double V = b->m.volume[e];
    double pl = 0.0, pm = 0.0;
    for (int i = 0; i < 9; ++i) {
      pl += (J - 1.0) * C[i] * D[i];
      pm += ((F[i] - C[i]) + (J - 1.0) * C[i]) * D[i];
    }
    lam_bar[e] -= V * pl;
    mu_bar[e] -= V * pm;
    if (X_bar) {
      /* dPhi/dDm = V[(P:D) I - F^T dP(D) - D^T P] Dm^-T, then the mass term
       * (rho/4) sum_{i in e} w_i V Dm^-T */
      snh_P(F, b->lambda[e], b->mu[e], P);
      snh_dP(F, D, b->lambda[e], b->mu[e], dP);
      const double PD = contract3(P, D);
      const uint32_t *t = &b->m.tet[4 * e];
      const double we = w[t[0]] + w[t[1]] + w[t[2]] + w[t[3]];
      double A[9], G[9];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double fq = 0.0, dp = 0.0;
          for (int k = 0; k < 3; ++k) {
            fq += F[k * 3 + r] * dP[k * 3 + c];
            dp += D[k * 3 + r] * P[k * 3 + c];
          }
          A[r * 3 + c] = V * ((r == c ? PD : 0.0) - fq - dp) + (r == c ? 0.25 * b->rho[e] * we * V : 0.0);
        }
      mul3_nt(A, dm, G); /* G = A Dm^-T */
      for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r) {
          X_bar[3 * t[k + 1] + r] -= G[r * 3 + k];
          X_bar[3 * t[0] + r] += G[r * 3 + k];
        }
    }
  }
  for (size_t n = 0; n < N; ++n)
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      const double x1 = xb[i], v1 = vb[i];
      if (b->m.anchored[n]) {
        xb[i] = x1 - Kl[i];
        vb[i] = -(1.0 - rule->af) * inner * Kl[i];
        ab[i] = -(1.0 - rule->af) * carry * Kl[i];
      } else {
        xb[i] = x1 - Kl[i];
        vb[i] = inner * x1 + v1 - (1.0 - rule->af) * inner * Kl[i];
        ab[i] = carry * x1 + inner * (1.0 - rule->gamma) * v1 - rule->am * b->m.mass[n] * lam[i] -
                (1.0 - rule->af) * carry * Kl[i];
      }
    }
  free(xw);
  free(seed);
  free(lam);
  free(Kl);
  free(w);
}

int orc_rollout_backward(orc_body *b, const double *x_bar, const double *v_bar, double *lam_bar,
                         double *mu_bar, double *X_bar, double *x0_bar, double *v0_bar,
                         double *a0_bar, double tolerance) {
  const size_t N = b->m.nodes, T = b->m.tets;
  if (!b->ckpt) return -1;
  for (size_t e = 0; e < T; ++e)
    if (!b->snh[e] || b->plastic_lane[e]) return -2;
  const orc_settings *s = &b->ckpt_settings;
  const rule_t rule = rule_of(s->spectral_radius);
  const uint32_t configured = s->substeps > 1 ? s->substeps : 1;
  double *xb = (double *)malloc(3 * N * sizeof(double));
  double *vb = (double *)malloc(3 * N * sizeof(double));
  double *ab = (double *)calloc(3 * N, sizeof(double));
  memcpy(xb, x_bar, 3 * N * sizeof(double));
  if (v_bar) memcpy(vb, v_bar, 3 * N * sizeof(double));
  else memset(vb, 0, 3 * N * sizeof(double));
  memset(lam_bar, 0, T * sizeof(double));
  memset(mu_bar, 0, T * sizeof(double));
  if (X_bar) memset(X_bar, 0, 3 * N * sizeof(double));
  const uint32_t every = b->ckpt_every, steps = b->ckpt_steps;
  const uint32_t segments = (steps + every - 1) / every;
  /* per substep of a segment: x_n, v_n, a_n, a', inner */
  const size_t per = 12 * N;
  for (uint32_t seg = segments; seg-- > 0;) {
    const double *c = &b->ckpt[(size_t)seg * 9 * N];
    memcpy(b->x, c, 3 * N * sizeof(double));
    memcpy(b->v, c + 3 * N, 3 * N * sizeof(double));
    memcpy(b->a, c + 6 * N, 3 * N * sizeof(double));
    const uint32_t t0 = seg * every, t1 = t0 + every < steps ? t0 + every : steps;
    size_t cap = 0, used = 0;
    double *rec = NULL, *inner_of = NULL;
    for (uint32_t t = t0; t < t1; ++t) {
      uint32_t substeps = configured;
      if (predictor_inverts(b, &rule, (double)s->dt / configured))
        substeps = configured > 4 ? configured : 4;
      const double inner = (double)s->dt / substeps;
      for (uint32_t k = 0; k < substeps; ++k) {
        if (used == cap) {
          cap = cap ? 2 * cap : 16;
          rec = (double *)realloc(rec, cap * per * sizeof(double));
          inner_of = (double *)realloc(inner_of, cap * sizeof(double));
        }
        double *r = &rec[used * per];
        memcpy(r, b->x, 3 * N * sizeof(double));
        memcpy(r + 3 * N, b->v, 3 * N * sizeof(double));
        memcpy(r + 6 * N, b->a, 3 * N * sizeof(double));
        orc_report rep;
        memset(&rep, 0, sizeof rep);
        substep(b, s, &rule, inner, &rep);
        memcpy(r + 9 * N, b->a, 3 * N * sizeof(double));
        inner_of[used++] = inner;
      }
    }
    for (size_t k = used; k-- > 0;) {
      const double *r = &rec[k * per];
      adjoint_substep(b, &rule, inner_of[k], s->gravity, r, r + 3 * N, r + 6 * N, r + 9 * N, xb,
                      vb, ab, lam_bar, mu_bar, X_bar, tolerance);
    }
    free(rec);
    free(inner_of);
  }
  if (x0_bar) memcpy(x0_bar, xb, 3 * N * sizeof(double));
  if (v0_bar) memcpy(v0_bar, vb, 3 * N * sizeof(double));
  if (a0_bar) memcpy(a0_bar, ab, 3 * N * sizeof(double));
  memcpy(b->x, b->final_state, 3 * N * sizeof(double));
  memcpy(b->v, b->final_state + 3 * N, 3 * N * sizeof(double));
  memcpy(b->a, b->final_state + 6 * N, 3 * N * sizeof(double));
  free(xb);
  free(vb);
  free(ab);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Boundary surface of a tet mesh (collectSurface's tet faces,          */
/* solver.cc:220-364) and crack faces (markCracks, tearing.cc:406-461,  */
/* restated for tets).                                                  */
/* ------------------------------------------------------------------ */

typedef struct {
  uint32_t s[3]; /* sorted nodes */
  uint32_t tet;
  uint32_t face;
} orc_face_t;

static int face_cmp(const void *pa, const void *pb) {
  const orc_face_t *a = (const orc_face_t *)pa, *b = (const orc_face_t *)pb;
  for (int k = 0; k < 3; ++k)
    if (a->s[k] != b->s[k]) return a->s[k] < b->s[k] ? -1 : 1;
  if (a->tet != b->tet) return a->tet < b->tet ? -1 : 1;
  return a->face < b->face ? -1 : (a->face > b->face);
}

typedef struct {
  double p[9];
  uint32_t idx;
} orc_spot_t;

static int spot_cmp(const void *pa, const void *pb) {
  const orc_spot_t *a = (const orc_spot_t *)pa, *b = (const orc_spot_t *)pb;
  for (int k = 0; k < 9; ++k)
    if (a->p[k] != b->p[k]) return a->p[k] < b->p[k] ? -1 : 1;
  return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

static int spot_cmp_points(const orc_spot_t *a, const orc_spot_t *b) {
  for (int k = 0; k < 9; ++k)
    if (a->p[k] != b->p[k]) return a->p[k] < b->p[k] ? -1 : 1;
  return 0;
}

static int point_cmp(const void *pa, const void *pb) {
  const double *a = (const double *)pa, *b = (const double *)pb;
  for (int k = 0; k < 3; ++k)
    if (a[k] != b[k]) return a[k] < b[k] ? -1 : 1;
  return 0;
}

static int edge_cmp(const void *pa, const void *pb) {
  const uint32_t *a = (const uint32_t *)pa, *b = (const uint32_t *)pb;
  if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
  return a[1] < b[1] ? -1 : (a[1] > b[1]);
}

/* Outputs: tri (3 per surface face, capacity 3*4*tets), tri_tet, tri_lane
 * (4*tets), edges (2 per edge, capacity 2*3*4*tets), surface nodes
 * (capacity nodes), crack (tets).  Returns the triangle count; *edges_out and
 * *nodes_out receive the edge and node counts. */
uint32_t orc_surface(uint32_t nodes, uint32_t tets, const uint32_t *tet, const uint32_t *lane,
                     const double *rest, uint32_t *tri, uint32_t *tri_tet, uint32_t *tri_lane,
                     uint32_t *edges, uint32_t *edges_out, uint32_t *surf, uint32_t *nodes_out,
                     uint8_t *crack) {
  static const int corner[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  const size_t nf = 4 * (size_t)tets;
  orc_face_t *faces = (orc_face_t *)malloc((nf ? nf : 1) * sizeof *faces);
  for (size_t e = 0; e < tets; ++e)
    for (int f = 0; f < 4; ++f) {
      orc_face_t *q = &faces[4 * e + f];
      uint32_t a = tet[4 * e + corner[f][0]], b = tet[4 * e + corner[f][1]], c = tet[4 * e + corner[f][2]], t;
      if (a > b) { t = a; a = b; b = t; }
      if (b > c) { t = b; b = c; c = t; }
      if (a > b) { t = a; a = b; b = t; }
      q->s[0] = a; q->s[1] = b; q->s[2] = c;
      q->tet = (uint32_t)e;
      q->face = (uint32_t)f;
    }
  qsort(faces, nf, sizeof *faces, face_cmp);
  uint32_t nt = 0, ne = 0;
  uint8_t *on = (uint8_t *)calloc(nodes ? nodes : 1, 1);
  memset(crack, 0, tets);
  orc_spot_t *spots = (orc_spot_t *)malloc((nf ? nf : 1) * sizeof *spots);
  for (size_t i = 0; i < nf;) {
    size_t j = i + 1;
    while (j < nf && faces[j].s[0] == faces[i].s[0] && faces[j].s[1] == faces[i].s[1] &&
           faces[j].s[2] == faces[i].s[2])
      ++j;
    if (j - i == 1) {
      const orc_face_t *q = &faces[i];
      const uint32_t *tn = &tet[4 * (size_t)q->tet];
      const int *c3 = corner[q->face];
      uint32_t a = tn[c3[0]], b = tn[c3[1]], c = tn[c3[2]];
      const uint32_t d = tn[6 - c3[0] - c3[1] - c3[2]];
      /* orientation: det[b-a, c-a, d-a] > 0 means the normal (b-a)x(c-a)
         points toward d, into the tet */
      double m[9];
      for (int r = 0; r < 3; ++r) {
        m[r * 3 + 0] = rest[3 * (size_t)b + r] - rest[3 * (size_t)a + r];
        m[r * 3 + 1] = rest[3 * (size_t)c + r] - rest[3 * (size_t)a + r];
        m[r * 3 + 2] = rest[3 * (size_t)d + r] - rest[3 * (size_t)a + r];
      }
      if (orc_det3(m) > 0.0) {
        const uint32_t t = b;
        b = c;
        c = t;
      }
      tri[3 * nt] = a; tri[3 * nt + 1] = b; tri[3 * nt + 2] = c;
      tri_tet[nt] = q->tet;
      tri_lane[nt] = lane ? lane[q->tet] : 0u;
      on[a] = on[b] = on[c] = 1;
      const uint32_t pairs[3][2] = {{q->s[0], q->s[1]}, {q->s[0], q->s[2]}, {q->s[1], q->s[2]}};
      for (int k = 0; k < 3; ++k) {
        edges[2 * ne] = pairs[k][0];
        edges[2 * ne + 1] = pairs[k][1];
        ++ne;
      }
      orc_spot_t *sp = &spots[nt];
      for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r) sp->p[3 * k + r] = rest[3 * (size_t)q->s[k] + r];
      qsort(sp->p, 3, 3 * sizeof(double), point_cmp);
      sp->idx = nt;
      ++nt;
    }
    i = j;
  }
  qsort(edges, ne, 2 * sizeof(uint32_t), edge_cmp);
  uint32_t unique = 0;
  for (uint32_t k = 0; k < ne; ++k)
    if (!unique || edges[2 * k] != edges[2 * (unique - 1)] || edges[2 * k + 1] != edges[2 * (unique - 1) + 1]) {
      edges[2 * unique] = edges[2 * k];
      edges[2 * unique + 1] = edges[2 * k + 1];
      ++unique;
    }
  *edges_out = unique;
  uint32_t ns = 0;
  for (uint32_t v = 0; v < nodes; ++v)
    if (on[v]) surf[ns++] = v;
  *nodes_out = ns;
  qsort(spots, nt, sizeof *spots, spot_cmp);
  for (uint32_t i = 0; i < nt;) {
    uint32_t j = i + 1;
    while (j < nt && spot_cmp_points(&spots[j], &spots[i]) == 0) ++j;
    if (j - i > 1)
      for (uint32_t k = i; k < j; ++k) {
        const uint32_t tr = spots[k].idx;
        /* the face index of triangle tr: find it among its tet's faces */
        const uint32_t e = tri_tet[tr];
        for (int f = 0; f < 4; ++f) {
          uint32_t s3[3] = {tet[4 * (size_t)e + corner[f][0]], tet[4 * (size_t)e + corner[f][1]],
                            tet[4 * (size_t)e + corner[f][2]]};
          int same = 1;
          for (int r = 0; r < 3 && same; ++r)
            same = s3[r] == tri[3 * tr + 0] || s3[r] == tri[3 * tr + 1] || s3[r] == tri[3 * tr + 2];
          if (same) crack[e] = (uint8_t)(crack[e] | (1u << f));
        }
      }
    i = j;
  }
  free(spots);
  free(on);
  free(faces);
  return nt;
}

/* clusterSurface (contact.cc:540-597; spreadBits 483-491: 10 bits to every
 * third bit) */
static uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

static int cmp_keyed(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

uint32_t orc_cluster_surface(const uint32_t *tri, uint32_t ntri, const double *rest, uint32_t *faces,
                             uint32_t *starts) {
  if (ntri == 0) return 0;
  float *centre = (float *)malloc(3 * (size_t)ntri * sizeof(float));
  uint64_t *keyed = (uint64_t *)malloc((size_t)ntri * sizeof(uint64_t));
  float low[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, high[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (uint32_t f = 0; f < ntri; ++f)
    for (int k = 0; k < 3; ++k) {
      const float sum = (float)rest[3 * (size_t)tri[3 * f] + k] + (float)rest[3 * (size_t)tri[3 * f + 1] + k] +
                        (float)rest[3 * (size_t)tri[3 * f + 2] + k];
      const float c = sum / 3.0f;
      centre[3 * f + k] = c;
      low[k] = c < low[k] ? c : low[k];
      high[k] = c > high[k] ? c : high[k];
    }
  float span[3];
  for (int k = 0; k < 3; ++k) span[k] = high[k] - low[k] > FLT_MIN ? high[k] - low[k] : FLT_MIN;
  for (uint32_t f = 0; f < ntri; ++f) {
    uint32_t cell[3];
    for (int k = 0; k < 3; ++k) {
      float t = (centre[3 * f + k] - low[k]) / span[k];
      t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
      cell[k] = (uint32_t)(t * 1023.0f);
    }
    const uint32_t key = spread10(cell[0]) | (spread10(cell[1]) << 1) | (spread10(cell[2]) << 2);
    keyed[f] = ((uint64_t)key << 32) | f; /* key, then face */
  }
  qsort(keyed, ntri, sizeof(uint64_t), cmp_keyed);
  for (uint32_t f = 0; f < ntri; ++f) faces[f] = (uint32_t)(keyed[f] & 0xffffffffu);
  uint32_t c = 0;
  for (uint32_t s = 0; s < ntri; s += 32) starts[c++] = s;
  starts[c] = ntri;
  free(centre);
  free(keyed);
  return c;
}
