// tmath.h -- per-tet constitutive math for gfx950 kernels and host code.
//
// Templated on the element precision Real (double for the plain FP64 path,
// float for the mixed path).  3x3 matrices are row-major arrays held in
// registers.  Reference formulas (paths relative to the Talus tree):
//   shape gradients   src/talus/deformable/material/material.cc:199-207
//   polar rotation    material.cc:168-197 (24 passes, tolerance 1e-6)
//   inverted frame    material.cc:24-64 (sign-fixed SVD when det F <= 0)
//   SNH stress        src/talus/dynamics/neohookean/neohookean.cc:41-47,
//                     125-147; material.cc:918-946 (displacement form)
//   J2 return map     material.cc:594-630
//   Hencky return     src/talus/dynamics/strainstep/plasticflow.cc:110-255
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define TM_HD __host__ __device__ __forceinline__

namespace tmi {

template <class R>
TM_HD R det3(const R *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// cofactor matrix, cof(A) = det(A) A^-T
template <class R>
TM_HD void cof3(const R *a, R *c) {
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

template <class R>
TM_HD bool inv3(const R *a, R *o) {
  R c[9];
  cof3(a, c);
  const R d = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
  if (d == R(0)) return false;
  const R s = R(1) / d;
  o[0] = c[0] * s; o[1] = c[3] * s; o[2] = c[6] * s;
  o[3] = c[1] * s; o[4] = c[4] * s; o[5] = c[7] * s;
  o[6] = c[2] * s; o[7] = c[5] * s; o[8] = c[8] * s;
  return true;
}

template <class R>
TM_HD void mul3(const R *a, const R *b, R *c) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      c[r * 3 + k] = a[r * 3 + 0] * b[k] + a[r * 3 + 1] * b[3 + k] + a[r * 3 + 2] * b[6 + k];
}

// c = a^T b
template <class R>
TM_HD void mul3_tn(const R *a, const R *b, R *c) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      c[r * 3 + k] = a[r] * b[k] + a[3 + r] * b[3 + k] + a[6 + r] * b[6 + k];
}

// c = a b^T
template <class R>
TM_HD void mul3_nt(const R *a, const R *b, R *c) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      c[r * 3 + k] = a[r * 3] * b[k * 3] + a[r * 3 + 1] * b[k * 3 + 1] + a[r * 3 + 2] * b[k * 3 + 2];
}

template <class R>
TM_HD void mv3(const R *a, const R *v, R *o) {
  const R x = v[0], y = v[1], z = v[2];
  o[0] = a[0] * x + a[1] * y + a[2] * z;
  o[1] = a[3] * x + a[4] * y + a[5] * z;
  o[2] = a[6] * x + a[7] * y + a[8] * z;
}

template <class R>
TM_HD void mtv3(const R *a, const R *v, R *o) {
  const R x = v[0], y = v[1], z = v[2];
  o[0] = a[0] * x + a[3] * y + a[6] * z;
  o[1] = a[1] * x + a[4] * y + a[7] * z;
  o[2] = a[2] * x + a[5] * y + a[8] * z;
}

template <class R>
TM_HD void ident3(R *a) {
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = (i % 4 == 0) ? R(1) : R(0);
}

// H = [u1-u0, u2-u0, u3-u0] Dm^-1 from displacements u (12) and Dm^-1 rows
template <class R>
TM_HD void displacement_gradient(const R *u, const R *dm, R *H) {
  R d[9];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int r = 0; r < 3; ++r) d[k * 3 + r] = u[(k + 1) * 3 + r] - u[r];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      H[r * 3 + c] = d[r] * dm[c] + d[3 + r] * dm[3 + c] + d[6 + r] * dm[6 + c];
}

// gradients g0..g3 (12) from the rows of Dm^-1
template <class R>
TM_HD void shape_gradients(const R *dm, R *g) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    g[3 + c] = dm[c];
    g[6 + c] = dm[3 + c];
    g[9 + c] = dm[6 + c];
    g[c] = -(dm[c] + dm[3 + c] + dm[6 + c]);
  }
}

// J - 1 = tr H + (tr^2 H - H:H^T)/2 + det H   (volumeChange)
template <class R>
TM_HD R volume_change(const R *H) {
  const R tr = H[0] + H[4] + H[8];
  R ht = R(0);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) ht += H[r * 3 + c] * H[c * 3 + r];
  return tr + R(0.5) * (tr * tr - ht) + det3(H);
}

// Stable neo-Hookean P(H) and energy density (H-form, no polar needed).
template <class R>
TM_HD R snh_stress(const R *H, R lambda, R mu, R *P) {
  const R shifted = lambda + mu;
  const R tr = H[0] + H[4] + H[8];
  const R change = volume_change(H);
  R cofH[9], F[9], cofF[9];
  cof3(H, cofH);
#pragma unroll
  for (int i = 0; i < 9; ++i) F[i] = H[i];
  F[0] += R(1);
  F[4] += R(1);
  F[8] += R(1);
  cof3(F, cofF);
  R hh = R(0);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      P[r * 3 + c] = mu * (H[r * 3 + c] + H[c * 3 + r] - (r == c ? tr : R(0)) - cofH[r * 3 + c]) +
                     shifted * change * cofF[r * 3 + c];
      hh += H[r * 3 + c] * H[r * 3 + c];
    }
  return R(0.5) * mu * hh - mu * (change - tr) + R(0.5) * shifted * change * change;
}

// polarRotation: R <- (R + R^-T)/2 from R = F, <= 24 passes, stop when the
// largest entry change is below 1e-6; identity when |det F| < 1e-12.
template <class R>
TM_HD void polar(const R *F, R *out) {
  if (fabs(det3(F)) < R(1.0e-12)) {
    ident3(out);
    return;
  }
  R r[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) r[i] = F[i];
  for (int pass = 0; pass < 24; ++pass) {
    R c[9];
    cof3(r, c);
    const R d = r[0] * c[0] + r[1] * c[1] + r[2] * c[2];
    const R s = R(0.5) / d;
    R largest = R(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      const R next = R(0.5) * r[i] + c[i] * s;
      const R change = fabs(next - r[i]);
      largest = change > largest ? change : largest;
      r[i] = next;
    }
    if (largest < R(1.0e-6)) break;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i] = r[i];
}

// Two Newton steps of R <- (R + R^-T)/2 in FP64 (R^-T = cof(R)/det R): an
// FP32 polar factor (orthogonal to ~1e-7) becomes orthogonal to rounding.
TM_HD void orthonormalize(double *R) {
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    double c[9];
    cof3(R, c);
    const double inv = 1.0 / (R[0] * c[0] + R[1] * c[1] + R[2] * c[2]);
#pragma unroll
    for (int i = 0; i < 9; ++i) R[i] = 0.5 * (R[i] + c[i] * inv);
  }
}

// Symmetric 3x3 eigen decomposition by cyclic Jacobi in FP64; values
// ascending, eigenvector i in column i of V (row-major).
TM_HD void eigen_sym3(const double *A, double *w, double *V) {
  double a[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) a[i] = A[i];
  a[1] = a[3] = 0.5 * (A[1] + A[3]);
  a[2] = a[6] = 0.5 * (A[2] + A[6]);
  a[5] = a[7] = 0.5 * (A[5] + A[7]);
  double v[9];
  ident3(v);
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
    const double diag = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
    if (off <= 1e-36 * diag || off == 0.0) break;
    for (int k = 0; k < 3; ++k) {
      const int p = k == 2 ? 1 : 0, q = k == 0 ? 1 : 2;
      const double apq = a[p * 3 + q];
      if (apq == 0.0) continue;
      const double app = a[p * 3 + p], aqq = a[q * 3 + q];
      const double theta = (aqq - app) / (2.0 * apq);
      const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
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
  double d[3] = {a[0], a[4], a[8]};
  int o[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (d[o[j]] < d[o[i]]) {
        const int t = o[i];
        o[i] = o[j];
        o[j] = t;
      }
  for (int i = 0; i < 3; ++i) {
    w[i] = d[o[i]];
    for (int r = 0; r < 3; ++r) V[r * 3 + i] = v[r * 3 + o[i]];
  }
}

// Rotation of the sign-fixed SVD of F (det F <= 0): U diag(1,1,det) V^T with
// the reflection on the smallest singular value, built from the eigenvectors
// of F^T F in FP64.
TM_HD void svd_rotation(const double *F, double *R) {
  double C[9], w[3], V[9];
  mul3_tn(F, F, C);
  eigen_sym3(C, w, V);
  double v0[3], v1[3], v2[3];
  for (int r = 0; r < 3; ++r) {
    v0[r] = V[r * 3 + 2];
    v1[r] = V[r * 3 + 1];
  }
  v2[0] = v0[1] * v1[2] - v0[2] * v1[1];
  v2[1] = v0[2] * v1[0] - v0[0] * v1[2];
  v2[2] = v0[0] * v1[1] - v0[1] * v1[0];
  double u0[3], u1[3], u2[3];
  mv3(F, v0, u0);
  mv3(F, v1, u1);
  const double n0 = sqrt(u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
  if (!(n0 > 0.0)) {
    ident3(R);
    return;
  }
  for (int r = 0; r < 3; ++r) u0[r] /= n0;
  const double p = u1[0] * u0[0] + u1[1] * u0[1] + u1[2] * u0[2];
  for (int r = 0; r < 3; ++r) u1[r] -= p * u0[r];
  const double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  if (!(n1 > 0.0)) {
    ident3(R);
    return;
  }
  for (int r = 0; r < 3; ++r) u1[r] /= n1;
  u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
  u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
  u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R[r * 3 + c] = u0[r] * v0[c] + u1[r] * v1[c] + u2[r] * v2[c];
}

// J2 radial return (returnMap, material.cc:594-630) on the corotated small
// strain; plastic = {eps_p[9], alpha}.  Writes the stress, returns true when
// the point returns to the yield surface; return_mapping = false is the
// elastic trial.  lin, if given, receives the linearisation {theta,
// thetaBar, n00, n01, n02, n11, n12, n22} (theta 1, thetaBar 0 when elastic)
// and excess the relative overshoot (sigma_eq - radius) / radius.
template <class R>
TM_HD bool j2_return_lin(const R *strain, const double *plastic, R lambda, R mu, R yield, R hardening,
                         bool return_mapping, R *stress, double *next, R *lin, R *excess) {
  const R bulk = lambda + R(2) * mu / R(3);
  R el[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) el[i] = strain[i] - R(plastic[i]);
  const R tr = el[0] + el[4] + el[8];
  R trial[9];
  R nn = R(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    trial[i] = R(2) * mu * el[i] - ((i % 4 == 0) ? R(2) * mu * tr / R(3) : R(0));
    nn += trial[i] * trial[i];
  }
  const R norm = sqrt(nn);
  const R equivalent = R(1.2247448713915890491) * norm;
  const R radius = yield + hardening * R(plastic[9]);
  if (next) {
    for (int i = 0; i < 10; ++i) next[i] = plastic[i];
  }
  if (lin) {
    lin[0] = R(1);
#pragma unroll
    for (int i = 1; i < 8; ++i) lin[i] = R(0);
  }
  if (excess) *excess = R(0);
  if (!return_mapping || !(equivalent > radius) || !(norm > R(0))) {
#pragma unroll
    for (int i = 0; i < 9; ++i) stress[i] = trial[i] + ((i % 4 == 0) ? bulk * tr : R(0));
    return false;
  }
  const R gamma = (equivalent - radius) / (R(3) * mu + hardening);
  const R theta = R(1) - R(3) * mu * gamma / equivalent;
#pragma unroll
  for (int i = 0; i < 9; ++i) stress[i] = theta * trial[i] + ((i % 4 == 0) ? bulk * tr : R(0));
  if (next) {
    for (int i = 0; i < 9; ++i)
      next[i] += double(R(1.2247448713915890491) * gamma * trial[i] / norm);
    next[9] += double(gamma);
  }
  if (lin) {
    lin[0] = theta;
    lin[1] = R(3) * mu / (R(3) * mu + hardening) - (R(1) - theta);
    lin[2] = trial[0] / norm;
    lin[3] = trial[1] / norm;
    lin[4] = trial[2] / norm;
    lin[5] = trial[4] / norm;
    lin[6] = trial[5] / norm;
    lin[7] = trial[8] / norm;
  }
  if (excess) *excess = (equivalent - radius) / radius;
  return true;
}

template <class R>
TM_HD bool j2_return(const R *strain, const double *plastic, R lambda, R mu, R yield,
                     R hardening, R *stress, double *next) {
  return j2_return_lin(strain, plastic, lambda, mu, yield, hardening, true, stress, next, (R *)nullptr,
                       (R *)nullptr);
}

// tangentOf (material.cc:636-647) for the linearisation lin of j2_return_lin:
// the stress K tr(e) I + 2 mu theta dev(e) - 2 mu thetaBar (n:e) n of a
// symmetric strain e.
template <class R>
TM_HD void j2_tangent_stress(const R *lin, R lambda, R mu, const R *e, R *out) {
  const R bulk = lambda + R(2) * mu / R(3);
  const R tr = e[0] + e[4] + e[8];
  const R n[9] = {lin[2], lin[3], lin[4], lin[3], lin[5], lin[6], lin[4], lin[6], lin[7]};
  R along = R(0);
#pragma unroll
  for (int i = 0; i < 9; ++i) along += n[i] * e[i];
  const R shear = R(2) * mu * lin[0], flow = R(2) * mu * lin[1] * along;
#pragma unroll
  for (int i = 0; i < 9; ++i)
    out[i] = shear * (e[i] - ((i % 4 == 0) ? tr / R(3) : R(0))) + ((i % 4 == 0) ? bulk * tr : R(0)) - flow * n[i];
}

// [v]x, row-major: [v]x w = v x w
TM_HD void cross_matrix(const double *v, double *m) {
  m[0] = 0.0;
  m[1] = -v[2];
  m[2] = v[1];
  m[3] = v[2];
  m[4] = 0.0;
  m[5] = -v[0];
  m[6] = -v[1];
  m[7] = v[0];
  m[8] = 0.0;
}

// initialStress of a corotated tet (rotationDerivative, material.cc:
// 1077-1118, spreadOf 690-709), FP64: x, X corner positions now and at rest
// (12 each), g the shape gradients, R the element rotation, f its corner
// forces; twist_i = R (sum_j K_ij [o_j]x) R^T + handed [f_i]x,
// turn_i = spin [handed R g_i]x, spin = R (tr S I - S)^-1 R^T,
// S = sym(R^T F).  Returns false (no term) when F or tr S I - S is singular.
TM_HD bool initial_stress_tet(const double *x, const double *X, const double *g, double V, double lam,
                              double mu, const double *R, const double *f, double *twist, double *turn) {
  double F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) F[r * 3 + k] += x[3 * c + r] * g[3 * c + k];
  if (fabs(det3(F)) < 1.0e-12) return false;
  double RtF[9], spread[9], inv[9];
  mul3_tn(R, F, RtF);
  const double tr = RtF[0] + RtF[4] + RtF[8];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      spread[r * 3 + k] = (r == k ? tr : 0.0) - 0.5 * (RtF[r * 3 + k] + RtF[k * 3 + r]);
  if (!(det3(spread) > 0.0) || !inv3(spread, inv)) return false;
  double Ri[9], spin[9];
  mul3(R, inv, Ri);
  mul3_nt(Ri, R, spin);
  const double handed = det3(R) < 0.0 ? -1.0 : 1.0;
  double xc[3] = {0, 0, 0}, Xc[3] = {0, 0, 0};
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 3; ++r) {
      xc[r] += 0.25 * x[3 * c + r];
      Xc[r] += 0.25 * X[3 * c + r];
    }
  double o[12];
  for (int c = 0; c < 4; ++c) {
    double arm[3];
    for (int r = 0; r < 3; ++r) arm[r] = x[3 * c + r] - xc[r];
    mtv3(R, arm, &o[3 * c]);
    for (int r = 0; r < 3; ++r) o[3 * c + r] -= X[3 * c + r] - Xc[r];
  }
  for (int i = 0; i < 4; ++i) {
    double tw[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double *gi = &g[3 * i];
    for (int j = 0; j < 4; ++j) {
      const double *gj = &g[3 * j];
      const double gg = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
      double K[9], C[9], KC[9];
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) K[r * 3 + k] = V * (lam * gi[r] * gj[k] + mu * gj[r] * gi[k] + (r == k ? mu * gg : 0.0));
      cross_matrix(&o[3 * j], C);
      mul3(K, C, KC);
      for (int k = 0; k < 9; ++k) tw[k] += KC[k];
    }
    double Rt[9], RtR[9], Cf[9], rg[3], Cg[9];
    mul3(R, tw, Rt);
    mul3_nt(Rt, R, RtR);
    cross_matrix(&f[3 * i], Cf);
    for (int k = 0; k < 9; ++k) twist[9 * i + k] = RtR[k] + handed * Cf[k];
    mv3(R, gi, rg);
    for (int r = 0; r < 3; ++r) rg[r] *= handed;
    cross_matrix(rg, Cg);
    mul3(spin, Cg, &turn[9 * i]);
  }
  return true;
}

// applyInitialStress (material.cc:1157-1174): y_i += (twist_i turned +
// turn_i^T twisted) / 2 with turned = sum turn_k d_k, twisted = sum twist_k^T d_k
TM_HD void apply_initial_stress(const double *twist, const double *turn, const double *d, double *y) {
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

// Hencky return (plasticflow.cc returnMap): returns true and the returned
// elastic gradient when the trial stress exceeds `yield`.
TM_HD bool hencky_return(const double *elastic, double mu, double yield, double *returned,
                         double *trial_stress) {
  double b[9], w[3], V[9];
  mul3_nt(elastic, elastic, b);
  eigen_sym3(b, w, V);
  if (!(w[0] > 0.0)) return false;
  double dev[3];
  for (int i = 0; i < 3; ++i) dev[i] = 0.5 * log(w[i]);
  const double mean = (dev[0] + dev[1] + dev[2]) / 3.0;
  for (int i = 0; i < 3; ++i) dev[i] -= mean;
  const double size = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const double stress = 1.2247448713915890491 * 2.0 * mu * size;
  if (trial_stress) *trial_stress = stress;
  if (!(stress > yield) || size <= 1.0e-15) return false;
  double scale[3];
  for (int i = 0; i < 3; ++i) scale[i] = exp(dev[i] * (yield / stress - 1.0));
  double corr[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      corr[r * 3 + c] = V[r * 3] * scale[0] * V[c * 3] + V[r * 3 + 1] * scale[1] * V[c * 3 + 1] +
                        V[r * 3 + 2] * scale[2] * V[c * 3 + 2];
  mul3(corr, elastic, returned);
  return true;
}

// flowLattice for the single point of a tet; ep = Fp^-1 - I, alpha the
// hardening strain.  Returns true on a flow (yield) event.
TM_HD bool hencky_flow(const double *F, double mu, double yield, double hardening,
                       double plastic_max, double *ep, double *alpha) {
  if (!(det3(F) > 0.0)) return false;
  double plastic[9], elastic[9];
  for (int i = 0; i < 9; ++i) plastic[i] = ep[i] + ((i % 4 == 0) ? 1.0 : 0.0);
  mul3(F, plastic, elastic);
  const double start = yield + hardening * (*alpha);
  const double ratio = hardening / (3.0 * mu);
  double returned[9], after[9], sigma_trial = 0.0, s2 = 0.0;
  if (!hencky_return(elastic, mu, start, returned, &sigma_trial)) return false;
  const double yield_now = (start + ratio * sigma_trial) / (1.0 + ratio);
  if (!hencky_return(elastic, mu, yield_now, after, &s2)) return false;
  double Finv[9], next[9];
  if (!inv3(F, Finv)) return false;
  mul3(Finv, after, next);
  const double kept = det3(next);
  if (isfinite(plastic_max) && kept > 0.0) {
    double ninv[9];
    inv3(next, ninv);
    ninv[0] -= 1.0;
    ninv[4] -= 1.0;
    ninv[8] -= 1.0;
    double ss = 0.0;
    for (int i = 0; i < 9; ++i) ss += ninv[i] * ninv[i];
    const double size = sqrt(ss);
    if (size > plastic_max) {
      double capped[9], cinv[9];
      for (int i = 0; i < 9; ++i) capped[i] = (plastic_max / size) * ninv[i];
      capped[0] += 1.0;
      capped[4] += 1.0;
      capped[8] += 1.0;
      const double cv = det3(capped);
      if (!(cv > 1.0e-6)) return false;
      inv3(capped, cinv);
      const double s = cbrt(kept * cv);
      for (int i = 0; i < 9; ++i) next[i] = s * cinv[i];
    }
  }
  double change = 0.0;
  for (int i = 0; i < 9; ++i) change += fabs(next[i] - plastic[i]);
  if (!(change > 1.0e-12)) return false;
  for (int i = 0; i < 9; ++i) ep[i] = next[i] - ((i % 4 == 0) ? 1.0 : 0.0);
  *alpha += (sigma_trial - yield_now) / (3.0 * mu);
  return true;
}

// sqrt(lambda_max(F^T F)) and its eigenvector (tear stretch of a tet)
TM_HD double stretch_of(const double *F, double *direction) {
  double C[9], w[3], V[9];
  mul3_tn(F, F, C);
  eigen_sym3(C, w, V);
  if (direction) {
    const double n = sqrt(V[2] * V[2] + V[5] * V[5] + V[8] * V[8]);
    direction[0] = V[2] / n;
    direction[1] = V[5] / n;
    direction[2] = V[8] / n;
  }
  return sqrt(w[2] > 0.0 ? w[2] : 0.0);
}

// ------------------------------------------------------- stiffness damping
// A tet's damping force is tau P R K R^T P v (applyDamping, factor/factor.cc:
// 848-857): K the elastic stiffness of the rest gradients and Lame moduli,
// R the committed rotation, P the removal of the corners' rigid motion.
// The rigid fit of the corners (rigidFit, factor.cc:783-817): arms a_i from
// the centroid and J^-1, J = sum (|a_i|^2 I - a_i a_i^T); a nearly singular
// J (det J <= 1e-6 (tr J / 3)^3) takes the pseudo-inverse over eigenvalues
// above 1e-6 tr J / 3.  fit = {arms (12), J^-1 upper triangle xx xy xz yy
// yz zz (6)}.
constexpr int FIT_PLANES = 18;

TM_HD void rigid_fit(const double *corners, double *fit) {
  double c[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) c[r] += corners[3 * i + r] * 0.25;
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double *a = &fit[3 * i];
#pragma unroll
    for (int r = 0; r < 3; ++r) a[r] = corners[3 * i + r] - c[r];
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) J[r * 3 + k] += (r == k ? aa : 0.0) - a[r] * a[k];
  }
  const double third = (J[0] + J[4] + J[8]) / 3.0;
  double inv[9];
  if (!(det3(J) > 1.0e-6 * third * third * third) || !inv3(J, inv)) {
    double w[3], V[9];
    eigen_sym3(J, w, V);
#pragma unroll
    for (int k = 0; k < 9; ++k) inv[k] = 0.0;
    for (int k = 0; k < 3; ++k) {
      if (!(w[k] > 1.0e-6 * third)) continue;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q) inv[r * 3 + q] += V[r * 3 + k] * V[q * 3 + k] / w[k];
    }
  }
  fit[12] = inv[0];
  fit[13] = inv[1];
  fit[14] = inv[2];
  fit[15] = inv[4];
  fit[16] = inv[5];
  fit[17] = inv[8];
}

// removeRigid (factor.cc:819-846): v_i -= mean(v) + s x a_i with the spin
// s = J^-1 sum_i a_i x v_i
template <class R>
TM_HD void remove_rigid(const R *fit, R *v) {
  R mean[3] = {R(0), R(0), R(0)}, m[3] = {R(0), R(0), R(0)};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const R *a = &fit[3 * i], *x = &v[3 * i];
#pragma unroll
    for (int r = 0; r < 3; ++r) mean[r] += x[r] * R(0.25);
    m[0] += a[1] * x[2] - a[2] * x[1];
    m[1] += a[2] * x[0] - a[0] * x[2];
    m[2] += a[0] * x[1] - a[1] * x[0];
  }
  const R *J = &fit[12];
  const R s[3] = {J[0] * m[0] + J[1] * m[1] + J[2] * m[2], J[1] * m[0] + J[3] * m[1] + J[4] * m[2],
                  J[2] * m[0] + J[4] * m[1] + J[5] * m[2]};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const R *a = &fit[3 * i];
    v[3 * i + 0] -= mean[0] + (s[1] * a[2] - s[2] * a[1]);
    v[3 * i + 1] -= mean[1] + (s[2] * a[0] - s[0] * a[2]);
    v[3 * i + 2] -= mean[2] + (s[0] * a[1] - s[1] * a[0]);
  }
}

// out = P R K R^T P v for one tet: with D = sum_j (R^T (P v)_j) g_j^T,
// out_i = V R (mu (D + D^T) + lambda tr(D) I) g_i, then P again
template <class R>
TM_HD void damp_tet(const R *fit, const R *Rm, const R *g, R V, R lam, R mu, const R *v, R *out) {
  R rel[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) rel[k] = v[k];
  remove_rigid(fit, rel);
  R D[9] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    R dp[3];
    mtv3(Rm, &rel[3 * j], dp);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) D[r * 3 + k] += dp[r] * g[3 * j + k];
  }
  const R tr = D[0] + D[4] + D[8];
  R S[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) S[r * 3 + k] = V * (mu * (D[r * 3 + k] + D[k * 3 + r]) + (r == k ? lam * tr : R(0)));
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R t[3];
    mv3(S, &g[3 * i], t);
    mv3(Rm, t, &out[3 * i]);
  }
  remove_rigid(fit, out);
}

}  // namespace tmi
