// kernels.hip -- gfx950 kernels of the deformable tet step.
//
// Element kernels (one lane per tet, tile-mapped) evaluate the constitutive
// law and write per-(tet, corner) slots; node kernels (one lane per node)
// sum their incident slots in CSR order, so assembly is deterministic and
// atomic-free.  PCG keeps every scalar on the device: per-tile partial sums
// are reduced per body in a fixed order by each consuming block, so no host
// round trip is needed inside a solve.
//
// Reference mapping (paths relative to the Talus tree):
//   k_predict          solver.cc:3680-3701 (predictor, alpha-point weights)
//   k_predictor_inv    solver.cc:568-623 (predictorInverts)
//   k_assemble         solver.cc:1354-1470 (assembleElastic) with
//                      material.cc:978-1057 / neohookean.cc:125-196 per tet
//   k_balance          solver.cc:3783-3806 (balance), 1574-1593 (guard)
//   k_op_*             solver.cc:904-1048 (applySystem)
//   k_pcg_*            solver.cc:1904-2089 (conjugateGradient)
//   k_finish           solver.cc:4351-4368, 5829-5849
//   k_commit           solver.cc:1260-1346 (commitPlastic) and
//                      plasticflow.cc:152-260 (flowLattice, Hencky)
//   k_tear             tearing.cc:469-483 (overstretched)
#include <cfloat>

#include "kernels.h"
#include "tmath.h"

namespace tmi {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Fixed-order block sum; result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  const int waves = blockDim.x >> 6;
  for (int w = 0; w < waves; ++w) s += lds[w];
  return s;
}

// Body total of partial slot `slot` over the body's node tiles, summed in a
// fixed order (identical in every block of the body).
__device__ __forceinline__ double body_total(const double *partial, uint32_t t0, uint32_t t1,
                                             int slot, double *lds) {
  double s = 0.0;
  for (uint32_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) s += partial[t * NPART + slot];
  return block_sum(s, lds);
}

template <class Real>
__device__ __forceinline__ void load_dm(const Dev<Real> &d, uint32_t e, Real *dm) {
#pragma unroll
  for (int k = 0; k < 9; ++k) dm[k] = d.dminv_r[(size_t)k * d.T + e];
}

// Shape gradients of a plastic neo-Hookean tet's plastic reference,
// Fp^-T g_i = (I + ep)^T g_i; the rest gradients for every other tet.  The
// system operator's shear part is taken along them, so that it follows the
// elastic response P(F Fp^-1) Fp^-T after the tet has flowed.
template <class Real, class A = Real>
__device__ __forceinline__ void plastic_gradients(const Dev<Real> &d, uint32_t e, bool flowing,
                                                  const A *g, A *gs) {
#pragma unroll
  for (int i = 0; i < 12; ++i) gs[i] = g[i];
  if (!flowing) return;
  A ep[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) ep[k] = A(d.plastic[(size_t)k * d.T + e]);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      gs[3 * i + r] += ep[r] * g[3 * i] + ep[3 + r] * g[3 * i + 1] + ep[6 + r] * g[3 * i + 2];
}

template <class Real>
__device__ __forceinline__ void gather_vec(const Real *src, uint4 c, Real *out) {
  const uint32_t n[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    out[3 * i + 0] = src[3 * (size_t)n[i] + 0];
    out[3 * i + 1] = src[3 * (size_t)n[i] + 1];
    out[3 * i + 2] = src[3 * (size_t)n[i] + 2];
  }
}

__device__ __forceinline__ void gather_vec64(const double *src, uint4 c, double *out) {
  const uint32_t n[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    out[3 * i + 0] = src[3 * (size_t)n[i] + 0];
    out[3 * i + 1] = src[3 * (size_t)n[i] + 1];
    out[3 * i + 2] = src[3 * (size_t)n[i] + 2];
  }
}

// FP64 deformation gradient of tet e at FP64 positions `pos`.
template <class Real>
__device__ __forceinline__ void deformation64(const Dev<Real> &d, const double *pos, uint32_t e,
                                              uint4 c, double *F) {
  double x[12], X[12], u[12], dm[9];
  gather_vec64(pos, c, x);
  gather_vec64(d.X, c, X);
#pragma unroll
  for (int i = 0; i < 12; ++i) u[i] = x[i] - X[i];
#pragma unroll
  for (int k = 0; k < 9; ++k) dm[k] = d.dminv[(size_t)k * d.T + e];
  displacement_gradient(u, dm, F);
  F[0] += 1.0;
  F[4] += 1.0;
  F[8] += 1.0;
}

// corner c of a tet writes its 3 values at the CSR position pos.c of its
// slot, so each node's incident slots are one contiguous run
template <class Real>
__device__ __forceinline__ void store_slots(Real *slot, uint4 pos, const Real *f) {
  const uint32_t at[4] = {pos.x, pos.y, pos.z, pos.w};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    Real *o = slot + 3 * (size_t)at[c];
    o[0] = f[3 * c + 0];
    o[1] = f[3 * c + 1];
    o[2] = f[3 * c + 2];
  }
}

// ---------------------------------------------------------------- nodes

template <class Real>
__global__ __launch_bounds__(TILE) void k_predict(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
  const double af = P.af, inner = P.inner, carry = P.carry;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const size_t i = 3 * (size_t)n + r;
    const double a = d.a[i], x = d.x[i];
    const double shifted = x + inner * d.v[i] + carry * a;
    const double weighted = (1.0 - af) * shifted + af * x;
    d.carried[i] = a;
    d.shifted[i] = shifted;
    d.weighted[i] = weighted;
    d.u_eval[i] = weighted - d.X[i];
    d.x0[i] = Real(a);
  }
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_evaluate(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const size_t i = 3 * (size_t)n + r;
    d.u_eval[i] = d.weighted[i] + P.reach * d.a[i] - d.X[i];
  }
}

// Per-node sums over a node tile's slots.  The tile's nodes own one
// contiguous run of CSR positions [k0, k1); the workgroup streams it through
// LDS in coalesced chunks and each thread adds its node's part [s, e) of
// every chunk in ascending position order, the order of a plain per-node
// loop.  Every thread of the block calls it with the same k0, k1.
constexpr uint32_t RUN_CHUNK = 1024;  // slots per LDS chunk
template <class Real, int C>
__device__ __forceinline__ void tile_run_sum(const Real *slots, uint32_t k0, uint32_t k1, uint32_t s,
                                             uint32_t e, Real *acc, Real *buf) {
  for (uint32_t c0 = k0; c0 < k1; c0 += RUN_CHUNK) {
    const uint32_t c1 = min(k1, c0 + RUN_CHUNK);
    __syncthreads();
    const Real *src = slots + (size_t)C * c0;
    for (uint32_t idx = threadIdx.x; idx < uint32_t(C) * (c1 - c0); idx += TILE) buf[idx] = src[idx];
    __syncthreads();
    const uint32_t lo = max(s, c0), hi = min(e, c1);
    for (uint32_t kk = lo; kk < hi; ++kk)
#pragma unroll
      for (int r = 0; r < C; ++r) acc[r] += buf[C * (kk - c0) + r];
  }
}

template <class Real, bool DIAG, bool TAKEN>
__global__ __launch_bounds__(TILE) void k_balance(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  __shared__ Real buf[6 * RUN_CHUNK];
  const uint32_t n0 = d.nt_begin[t], n1 = d.nt_end[t];
  const uint32_t n = n0 + threadIdx.x;
  const bool own = n < n1 && !d.anchored[n];
  const uint32_t s = own ? d.inc_start[n] : 0u, e = own ? d.inc_start[n + 1] : 0u;
  Real f[3] = {Real(0), Real(0), Real(0)};
  Real D[6] = {Real(0), Real(0), Real(0), Real(0), Real(0), Real(0)};
  tile_run_sum<Real, 3>(d.slot_f, d.inc_start[n0], d.inc_start[n1], s, e, f, buf);
  if (DIAG) tile_run_sum<Real, 6>(d.slot_d, d.inc_start[n0], d.inc_start[n1], s, e, D, buf);
  if (n >= n1) return;
  Real *Pn = d.P + 6 * (size_t)n;
  if (d.anchored[n]) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      d.v[3 * (size_t)n + r] = 0.0;
      d.a[3 * (size_t)n + r] = 0.0;
      d.right[3 * (size_t)n + r] = 0.0;
    }
    if (DIAG) {
      Pn[0] = Real(1); Pn[1] = Real(0); Pn[2] = Real(0);
      Pn[3] = Real(1); Pn[4] = Real(0); Pn[5] = Real(1);
    }
    return;
  }
  const double m = d.mass[n];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const size_t i = 3 * (size_t)n + r;
    double rhs = double(f[r]) + m * P.g[r] - P.am * m * d.carried[i];
    if (TAKEN) rhs -= P.inertia * m * d.a[i];
    d.right[i] = rhs;
  }
  if (DIAG) {
    const Real reach = Real(P.reach), im = Real(P.inertia * m);
    Real A[9] = {reach * D[0] + im, reach * D[1], reach * D[2],
                 reach * D[1], reach * D[3] + im, reach * D[4],
                 reach * D[2], reach * D[4], reach * D[5] + im};
    Real inv[9];
    if (det3(A) > Real(0) && inv3(A, inv)) {
      Pn[0] = inv[0]; Pn[1] = inv[1]; Pn[2] = inv[2];
      Pn[3] = inv[4]; Pn[4] = inv[5]; Pn[5] = inv[8];
    } else {
      Pn[0] = Real(1); Pn[1] = Real(0); Pn[2] = Real(0);
      Pn[3] = Real(1); Pn[4] = Real(0); Pn[5] = Real(1);
    }
  }
}

template <class Real>
__device__ __forceinline__ void apply_P(const Real *Pn, const Real *r, Real *z) {
  z[0] = Pn[0] * r[0] + Pn[1] * r[1] + Pn[2] * r[2];
  z[1] = Pn[1] * r[0] + Pn[3] * r[1] + Pn[4] * r[2];
  z[2] = Pn[2] * r[0] + Pn[4] * r[1] + Pn[5] * r[2];
}

// r = right - A x0 (A x0 from the op slots when WITH_X0), z = P r, p = z;
// with FACTOR the preconditioned vector comes from the factor solve later
template <class Real, bool WITH_X0, bool FACTOR>
__global__ __launch_bounds__(TILE) void k_pcg_init(Dev<Real> d) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double rz = 0.0, rr = 0.0, bb = 0.0;
  if (n < d.nt_end[t]) {
    Real r[3] = {Real(0), Real(0), Real(0)}, z[3] = {Real(0), Real(0), Real(0)};
    if (!d.anchored[n]) {
      Real q[3] = {Real(0), Real(0), Real(0)};
      if (WITH_X0) {
        const uint32_t k1 = d.inc_start[n + 1];
        for (uint32_t k = d.inc_start[n]; k < k1; ++k) {
          q[0] += d.slot_f[3 * (size_t)k + 0];
          q[1] += d.slot_f[3 * (size_t)k + 1];
          q[2] += d.slot_f[3 * (size_t)k + 2];
        }
        const Real im = Real(P.inertia * d.mass[n]), reach = Real(P.reach);
#pragma unroll
        for (int c = 0; c < 3; ++c) q[c] = im * d.x0[3 * (size_t)n + c] + reach * q[c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double rhs = d.right[3 * (size_t)n + c];
        r[c] = Real(rhs - double(q[c]));
        bb += rhs * rhs;
      }
      if (!FACTOR) apply_P(d.P + 6 * (size_t)n, r, z);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t i = 3 * (size_t)n + c;
      d.r[i] = r[c];
      if (!FACTOR) {
        d.z[i] = z[c];
        d.p[i] = z[c];
      }
      d.dx[i] = Real(0);
      rz += double(r[c]) * double(z[c]);
      rr += double(r[c]) * double(r[c]);
    }
  }
  rz = block_sum(rz, lds);
  rr = block_sum(rr, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    d.partial[t * NPART + 0] = rz;
    d.partial[t * NPART + 1] = rr;
    d.partial[t * NPART + 2] = bb;
  }
}

// one block per body: starting scalars of the solve
template <class Real>
__global__ __launch_bounds__(TILE) void k_pcg_begin(Dev<Real> d) {
  __shared__ double lds[TILE / 64];
  const uint32_t b = blockIdx.x;
  const StepParams &P = d.params[b];
  const uint32_t t0 = d.body_nt0[b], t1 = d.body_nt1[b];
  const double rz = body_total(d.partial, t0, t1, 0, lds);
  const double rr = body_total(d.partial, t0, t1, 1, lds);
  const double bb = body_total(d.partial, t0, t1, 2, lds);
  if (threadIdx.x == 0) {
    double reference = sqrt(bb);
    if (!(reference > double(FLT_MIN))) reference = double(FLT_MIN);
    const double target = P.tolerance * reference;
    d.target2[b] = target * target;
    d.sq[b] = rz;
    d.done[b] = (!P.active || sqrt(rr) <= target) ? 1 : 0;
    d.iters[b] = 0;
    d.rr_last[b] = rr;
  }
}

// ---------------------------------------------------------------- tets

// Per-tet operator y_i = V R sigma(R^T d) g_i, sigma(D') = lambda tr(D') I +
// mu (D' + D'^T); equals sum_j R K_ij R^T d_j (applyStiffness).  Neo-Hookean
// tets keep the shear part and take the volumetric part along the volume
// frame W = cof(Fe) Fp^-T: y_i = V(R mu (D' + D'^T) g_i + lambda (W:D) W g_i)
// with D = sum_j d_j g_j^T, so W:D = sum_j d_j . (W g_j).  At rest both are
// the linear stiffness; under stretch the volumetric part follows the
// (lambda+mu)(J-1)^2 term's stiffening along cof(F).
// A = Real is the operator of the PCG; A = double (with Real = float) is
// the same operator in FP64 arithmetic with the FP64 rest data (Dm^-1, V,
// lambda, mu), the residual of FP64 iterative refinement.
template <class Real, class A = Real>
__device__ __forceinline__ void op_tet_values(const Dev<Real> &d, uint32_t e, const A *dv, A *y,
                                              bool tangent = false) {
  constexpr bool WIDE = sizeof(A) > sizeof(Real);
  A dm[9], g[12], gs[12], Rm[9];
  if constexpr (WIDE) {
#pragma unroll
    for (int k = 0; k < 9; ++k) dm[k] = d.dminv[(size_t)k * d.T + e];
  } else {
    load_dm(d, e, dm);
  }
  shape_gradients(dm, g);
  const uint8_t flags = d.flags[e];
  plastic_gradients(d, e, (flags & FLAG_SNH) && (flags & FLAG_PLASTIC), g, gs);
#pragma unroll
  for (int k = 0; k < 9; ++k) Rm[k] = A(d.R[(size_t)k * d.T + e]);
  A D[9] = {A(0), A(0), A(0), A(0), A(0), A(0), A(0), A(0), A(0)};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    A dp[3];
    mtv3(Rm, &dv[3 * j], dp);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) D[r * 3 + k] += dp[r] * gs[3 * j + k];
  }
  const bool snh = (flags & FLAG_SNH) != 0;
  const A lam = WIDE ? A(d.lam64[e]) : A(d.lam[e]), mu = WIDE ? A(d.mu64[e]) : A(d.mu[e]),
          V = WIDE ? A(d.vol64[e]) : A(d.vol[e]);
  const A tr = snh ? A(0) : D[0] + D[4] + D[8];
  A S[9];
  if (tangent && !snh && (flags & FLAG_PLASTIC)) {
    // the consistent J2 tangent (tangentOf) on the corotated strain sym(D')
    A lin[8], eps[9], sig[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) lin[k] = A(d.lin[(size_t)k * d.T + e]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) eps[r * 3 + k] = A(0.5) * (D[r * 3 + k] + D[k * 3 + r]);
    j2_tangent_stress(lin, lam, mu, eps, sig);
#pragma unroll
    for (int k = 0; k < 9; ++k) S[k] = V * sig[k];
  } else {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        S[r * 3 + k] = V * (mu * (D[r * 3 + k] + D[k * 3 + r]) + (r == k ? lam * tr : A(0)));
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    A tmp[3];
    mv3(S, &gs[3 * i], tmp);
    mv3(Rm, tmp, &y[3 * i]);
  }
  if (snh) {
    A Wm[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Wm[k] = A(d.W[(size_t)k * d.T + e]);
    A vol = A(0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      A wg[3];
      mv3(Wm, &g[3 * j], wg);
      vol += dv[3 * j] * wg[0] + dv[3 * j + 1] * wg[1] + dv[3 * j + 2] * wg[2];
    }
    vol *= V * lam;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      A wg[3];
      mv3(Wm, &g[3 * i], wg);
#pragma unroll
      for (int r = 0; r < 3; ++r) y[3 * i + r] += vol * wg[r];
    }
  }
}

template <class Real>
__device__ __forceinline__ void op_tet(const Dev<Real> &d, uint32_t e, const Real *dir, Real *y,
                                       bool tangent = false) {
  Real dv[12];
  gather_vec(dir, d.conn[e], dv);
  op_tet_values(d, e, dv, y, tangent);
}

// the damping term of a damped tet, tau C d scaled into the operator's
// slots: the gather multiplies the slots by reach, so they take
// (damp / reach) tau C d (applySystem's addDamping, solver.cc:1006-1007)
template <class Real>
__device__ __forceinline__ void op_damping(const Dev<Real> &d, uint32_t e, const Real *dv, Real scale,
                                           Real *y) {
  Real fit[FIT_PLANES], Rm[9], dm[9], g[12], out[12];
#pragma unroll
  for (int k = 0; k < FIT_PLANES; ++k) fit[k] = d.fit[(size_t)k * d.T + e];
#pragma unroll
  for (int k = 0; k < 9; ++k) Rm[k] = d.R[(size_t)k * d.T + e];
  load_dm(d, e, dm);
  shape_gradients(dm, g);
  damp_tet(fit, Rm, g, d.vol[e], d.lam[e], d.mu[e], dv, out);
  const Real w = scale * d.tau[e];
#pragma unroll
  for (int k = 0; k < 12; ++k) y[k] += w * out[k];
}

// The tile's node vectors src[3n..3n+2] for its node list, staged in LDS
// (3 tn_max values of dynamic shared memory): every node once per tile,
// read in ascending node order, instead of four scattered 12-byte gathers
// per tet.
template <class V>
__device__ __forceinline__ void stage_tile_nodes(const uint32_t *tn_start, const uint32_t *tn_node, uint32_t t,
                                                 const V *src, V *buf) {
  const uint32_t s0 = tn_start[t], ns = tn_start[t + 1] - s0;
  for (uint32_t i = threadIdx.x; i < ns; i += TILE) {
    const size_t n = tn_node[s0 + i];
    buf[3 * i + 0] = src[3 * n + 0];
    buf[3 * i + 1] = src[3 * n + 1];
    buf[3 * i + 2] = src[3 * n + 2];
  }
  __syncthreads();
}
template <class V>
__device__ __forceinline__ void gather_staged(const V *buf, uint2 cl, V *out) {
  const uint32_t n[4] = {cl.x & 0xffffu, cl.x >> 16, cl.y & 0xffffu, cl.y >> 16};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    out[3 * i + 0] = buf[3 * n[i] + 0];
    out[3 * i + 1] = buf[3 * n[i] + 1];
    out[3 * i + 2] = buf[3 * n[i] + 2];
  }
}

template <class Real, bool DAMP>
__global__ __launch_bounds__(TILE) void k_op_element(Dev<Real> d, const Real *dir, int parity) {
  extern __shared__ double staged_raw[];
  Real *staged = reinterpret_cast<Real *>(staged_raw);
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x), b = d.tt_body[t];
  if (!d.params[b].active || (parity >= 0 && d.done[parity * d.B + b])) return;
  stage_tile_nodes(d.tn_start, d.tn_node, t, dir, staged);
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  if (e >= d.tt_end[t]) return;
  Real dv[12], y[12];
  gather_staged(staged, d.conn_l[e], dv);
  op_tet_values(d, e, dv, y, d.params[b].tangent != 0u);
  if (DAMP && (d.flags[e] & FLAG_DAMP)) {
    const StepParams &P = d.params[b];
    op_damping(d, e, dv, Real(P.damp / P.reach), y);
  }
  store_slots(d.slot_f, d.slot_pos[e], y);
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_op_element_flat(Dev<Real> d, const Real *dir) {
  const uint32_t e = blockIdx.x * TILE + threadIdx.x;
  if (e >= d.T) return;
  Real y[12];
  op_tet(d, e, dir, y);
  store_slots(d.slot_f, d.slot_pos[e], y);
}

// q = inertia M p + h^2 sum(slots), anchored rows zeroed; partial p.q
template <class Real>
__global__ __launch_bounds__(TILE) void k_op_gather(Dev<Real> d, const Real *dir, Real *out,
                                                    int parity) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active || (parity >= 0 && d.done[parity * d.B + b])) return;
  __shared__ Real buf[3 * RUN_CHUNK];
  const uint32_t n0 = d.nt_begin[t], n1 = d.nt_end[t];
  const uint32_t n = n0 + threadIdx.x;
  const bool own = n < n1 && !d.anchored[n];
  Real q[3] = {Real(0), Real(0), Real(0)};
  tile_run_sum<Real, 3>(d.slot_f, d.inc_start[n0], d.inc_start[n1], own ? d.inc_start[n] : 0u,
                        own ? d.inc_start[n + 1] : 0u, q, buf);
  double pq = 0.0;
  if (n < n1) {
    if (own) {
      const Real im = Real(P.inertia * d.mass[n]), reach = Real(P.reach);
#pragma unroll
      for (int c = 0; c < 3; ++c) q[c] = im * dir[3 * (size_t)n + c] + reach * q[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      out[3 * (size_t)n + c] = q[c];
      pq += double(dir[3 * (size_t)n + c]) * double(q[c]);
    }
  }
  pq = block_sum(pq, lds);
  if (threadIdx.x == 0) d.partial[t * NPART + 0] = pq;
}

// rz partial after a factor solve (slot 1, or slot 0 and p = z at the start)
template <class Real, bool START>
__global__ __launch_bounds__(TILE) void k_pcg_rz(Dev<Real> d, int parity) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active || (!START && d.done[parity * d.B + b])) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double rz = 0.0;
  if (n < d.nt_end[t]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t i = 3 * (size_t)n + c;
      const Real z = d.z[i];
      if (START) d.p[i] = z;
      rz += double(d.r[i]) * double(z);
    }
  }
  rz = block_sum(rz, lds);
  if (threadIdx.x == 0) d.partial[t * NPART + (START ? 0 : 1)] = rz;
}

template <class Real, bool FACTOR>
__global__ __launch_bounds__(TILE) void k_pcg_update(Dev<Real> d, int parity) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active || d.done[parity * d.B + b]) return;
  const double curvature = body_total(d.partial, d.body_nt0[b], d.body_nt1[b], 0, lds);
  const double squared = d.sq[parity * d.B + b];
  if (!(curvature > 0.0) || !(squared > 0.0)) return;
  const Real stride = Real(squared / curvature);
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (n < d.nt_end[t]) {
    Real r[3], z[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t i = 3 * (size_t)n + c;
      d.dx[i] += stride * d.p[i];
      r[c] = d.r[i] - stride * d.q[i];
      d.r[i] = r[c];
    }
    if (!FACTOR) {
      if (d.anchored[n]) {
        z[0] = z[1] = z[2] = Real(0);
      } else {
        apply_P(d.P + 6 * (size_t)n, r, z);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (!FACTOR) {
        d.z[3 * (size_t)n + c] = z[c];
        rz += double(r[c]) * double(z[c]);
      }
      rr += double(r[c]) * double(r[c]);
    }
  }
  if (!FACTOR) rz = block_sum(rz, lds);
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0) {
    if (!FACTOR) d.partial[t * NPART + 1] = rz;
    d.partial[t * NPART + 2] = rr;
  }
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_pcg_direction(Dev<Real> d, int parity) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active) return;
  const bool first = t == d.body_nt0[b];
  const int next_parity = parity ^ 1;
  if (d.done[parity * d.B + b]) {
    if (first && threadIdx.x == 0) d.done[next_parity * d.B + b] = 1;
    return;
  }
  const uint32_t t0 = d.body_nt0[b], t1 = d.body_nt1[b];
  const double curvature = body_total(d.partial, t0, t1, 0, lds);
  const double squared = d.sq[parity * d.B + b];
  if (!(curvature > 0.0) || !(squared > 0.0)) {
    if (first && threadIdx.x == 0) d.done[next_parity * d.B + b] = 1;
    return;
  }
  const double rz = body_total(d.partial, t0, t1, 1, lds);
  const double rr = body_total(d.partial, t0, t1, 2, lds);
  const Real blend = Real(rz / squared);
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n < d.nt_end[t]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t i = 3 * (size_t)n + c;
      d.p[i] = d.z[i] + blend * d.p[i];
    }
  }
  if (first && threadIdx.x == 0) {
    d.sq[next_parity * d.B + b] = rz;
    d.iters[b] += 1;
    d.rr_last[b] = rr;
    d.done[next_parity * d.B + b] = rr <= d.target2[b] ? 1 : 0;
  }
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_apply_increment(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t] || d.anchored[n]) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) d.a[3 * (size_t)n + c] += double(d.dx[3 * (size_t)n + c]);
}

// FP64 iterative refinement.  The corner values of the operator applied to
// the acceleration, in FP64 arithmetic (op_tet_values<Real, double>), into
// the FP64 slots.
// The unknown of the solve just made: the acceleration a (a solve from the
// previous acceleration), or the FP64 sum inc of the increments a
// correction solve and its refinements added to a (not a minus its old
// value, which would cancel most of the increment's digits).
__device__ __forceinline__ double unknown(const double *a, const double *inc, size_t i) {
  return inc ? inc[i] : a[i];
}

// inc = dx (add = false) or inc += dx: the FP64 sum of a correction solve's
// increments, exact sums of the FP32 (or FP64) increments up to FP64 rounding
template <class Real>
__global__ __launch_bounds__(TILE) void k_track_increment(Dev<Real> d, int add) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = 3 * (size_t)n + c;
    const double dx = d.anchored[n] ? 0.0 : double(d.dx[i]);
    d.rbase[i] = add ? d.rbase[i] + dx : dx;
  }
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_op_element64(Dev<Real> d, const double *inc) {
  const uint32_t t = blockIdx.x, b = d.tt_body[t];
  if (!d.params[b].active) return;
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  if (e >= d.tt_end[t]) return;
  double dv[12], y[12];
  const uint4 c4 = d.conn[e];
  const uint32_t nn[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) dv[3 * i + r] = unknown(d.a, inc, 3 * (size_t)nn[i] + r);
  op_tet_values<Real, double>(d, e, dv, y, d.params[b].tangent != 0u);
  if (d.fit && (d.flags[e] & FLAG_DAMP)) {
    // the damping term as the PCG's operator adds it (op_damping), in FP64
    const StepParams &P = d.params[b];
    double fit[FIT_PLANES], Rm[9], dm[9], g[12], out[12];
#pragma unroll
    for (int k = 0; k < FIT_PLANES; ++k) fit[k] = double(d.fit[(size_t)k * d.T + e]);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      Rm[k] = double(d.R[(size_t)k * d.T + e]);
      dm[k] = d.dminv[(size_t)k * d.T + e];
    }
    shape_gradients(dm, g);
    damp_tet(fit, Rm, g, d.vol64[e], d.lam64[e], d.mu64[e], dv, out);
    const double w = (P.damp / P.reach) * double(d.tau[e]);
#pragma unroll
    for (int k = 0; k < 12; ++k) y[k] += w * out[k];
  }
  const uint4 p = d.slot_pos[e];
  const uint32_t pos[4] = {p.x, p.y, p.z, p.w};
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) d.slot64[3 * (size_t)pos[c] + r] = y[3 * c + r];
}

// rres = right - (inertia M a + reach sum of the node's slots), anchored
// rows 0; per node tile |rres|^2 and |right|^2
template <class Real>
__global__ __launch_bounds__(TILE) void k_residual64(Dev<Real> d, const double *inc) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (n < d.nt_end[t]) {
    double q[3] = {0.0, 0.0, 0.0};
    if (!d.anchored[n]) {
      for (uint32_t k = d.inc_start[n]; k < d.inc_start[n + 1]; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) q[c] += d.slot64[3 * (size_t)k + c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t i = 3 * (size_t)n + c;
      double r = 0.0;
      if (!d.anchored[n]) {
        const double right = d.right[i];
        r = right - (P.inertia * d.mass[n] * unknown(d.a, inc, i) + P.reach * q[c]);
        bb += right * right;
      }
      d.rres[i] = r;
      rr += r * r;
    }
  }
  rr = block_sum(rr, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    d.rpart[2 * t] = rr;
    d.rpart[2 * t + 1] = bb;
  }
}

// one block per body: the relative FP64 residual, the refinement test, and
// a zero right-hand side for the bodies that meet it
template <class Real>
__global__ __launch_bounds__(TILE) void k_refine_check(Dev<Real> d, double tolerance) {
  __shared__ double lds[TILE / 64];
  __shared__ int met;
  const uint32_t b = blockIdx.x;
  if (!d.params[b].active) {
    if (threadIdx.x == 0) {
      d.rdone[b] = 1;
      d.rrel[b] = 0.0;
    }
    return;
  }
  const uint32_t t0 = d.body_nt0[b], t1 = d.body_nt1[b];
  double rr = 0.0, bb = 0.0;
  for (uint32_t t = t0 + threadIdx.x; t < t1; t += TILE) {
    rr += d.rpart[2 * t];
    bb += d.rpart[2 * t + 1];
  }
  rr = block_sum(rr, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    const double rel = bb > 0.0 ? sqrt(rr / bb) : (rr > 0.0 ? INFINITY : 0.0);
    d.rrel[b] = rel;
    met = rel <= tolerance ? 1 : 0;
    d.rdone[b] = met;
  }
  __syncthreads();
  if (!met) return;
  const uint32_t n0 = d.nt_begin[t0], n1 = d.nt_end[t1 - 1];
  for (uint32_t i = 3 * n0 + threadIdx.x; i < 3 * n1; i += TILE) d.rres[i] = 0.0;
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_finish(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = 3 * (size_t)n + c;
    if (d.anchored[n]) {
      d.v[i] = 0.0;
      d.a[i] = 0.0;
      continue;
    }
    const double a = d.a[i];
    d.v[i] += P.inner * ((1.0 - P.gamma) * d.carried[i] + P.gamma * a);
    d.x[i] = d.shifted[i] + P.beta * P.inner * P.inner * a;
  }
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_kinetic(Dev<Real> d) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double k = 0.0, bad = 0.0;
  if (n < d.nt_end[t]) {
    double vv = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double v = d.v[3 * (size_t)n + c];
      vv += v * v;
      if (!isfinite(v) || !isfinite(d.x[3 * (size_t)n + c])) bad = 1.0;
    }
    k = 0.5 * d.mass[n] * vv;
  }
  k = block_sum(k, lds);
  bad = block_sum(bad, lds);
  if (threadIdx.x == 0) {
    d.partial[t * NPART + 3] = k;
    d.partial[t * NPART + 1] = bad;
  }
}

// ------------------------------------------------------- element assembly

// Stiffness damping of one tet at its assembly (DAMP instantiations only):
// the rigid fit at the evaluation point (fitDamping, solver.cc:678-714),
// the viscous force tau C v_w at the alpha-point velocity v_w = v +
// lag((1 - gamma) a_n [+ gamma a]) (anchored corners at rest) subtracted
// from the corner forces (balance, solver.cc:3746-3783), and with DIAG the
// damping diagonal blocks tau C_ii (blockDamping / addDampingDiagonal,
// solver.cc:802-880) scaled by damp / reach, so that the balance's
// reach sum(slot_d) carries reach K_ii + damp tau C_ii.
template <class Real>
__device__ __forceinline__ void damp_assembly(const Dev<Real> &d, const StepParams &P, uint32_t e, uint4 c,
                                              const double *u64, const Real *g, const Real *Rm, Real V,
                                              Real lam, Real mu, int viscous, Real *f, Real *fit) {
  const Real tau = d.tau[e];
  double xc[12], fit64[FIT_PLANES];
  gather_vec64(d.X, c, xc);
#pragma unroll
  for (int k = 0; k < 12; ++k) xc[k] += u64[k];
  rigid_fit(xc, fit64);
#pragma unroll
  for (int k = 0; k < FIT_PLANES; ++k) {
    fit[k] = Real(fit64[k]);
    d.fit[(size_t)k * d.T + e] = fit[k];
  }
  if (viscous != VISCOUS_NONE) {
    const uint32_t nn[4] = {c.x, c.y, c.z, c.w};
    const double lc = P.lag * (1.0 - P.gamma), lt = viscous == VISCOUS_TAKEN ? P.lag * P.gamma : 0.0;
    Real vw[12], out[12];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool held = d.anchored[nn[i]] != 0;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const size_t k = 3 * (size_t)nn[i] + r;
        vw[3 * i + r] = held ? Real(0) : Real(d.v[k] + lc * d.carried[k] + lt * d.a[k]);
      }
    }
    damp_tet(fit, Rm, g, V, lam, mu, vw, out);
#pragma unroll
    for (int k = 0; k < 12; ++k) f[k] -= tau * out[k];
  }
}

template <class Real>
__device__ __forceinline__ void damp_diagonal(const Dev<Real> &d, const StepParams &P, uint32_t e,
                                              const Real *fit, const Real *g, const Real *Rm, Real V,
                                              Real lam, Real mu) {
  {
    const Real w = d.tau[e] * Real(P.damp / P.reach);
    const uint4 pos = d.slot_pos[e];
    const uint32_t at[4] = {pos.x, pos.y, pos.z, pos.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      Real B[9];
#pragma unroll
      for (int axis = 0; axis < 3; ++axis) {
        Real unit[12], col[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) unit[k] = Real(k == 3 * i + axis ? 1 : 0);
        damp_tet(fit, Rm, g, V, lam, mu, unit, col);
#pragma unroll
        for (int r = 0; r < 3; ++r) B[r * 3 + axis] = col[3 * i + r];
      }
      Real *o = d.slot_d + 6 * (size_t)at[i];
      o[0] += w * B[0];
      o[1] += w * B[1];
      o[2] += w * B[2];
      o[3] += w * B[4];
      o[4] += w * B[5];
      o[5] += w * B[8];
    }
  }
}

template <class Real, bool DIAG, bool ENERGY, bool DAMP>
__global__ __launch_bounds__(TILE) void k_assemble(Dev<Real> d, int viscous) {
  constexpr bool MIXED = sizeof(Real) < 8;
  __shared__ double lds[TILE / 64];
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x), b = d.tt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  double psi_total = 0.0;
  if (e < d.tt_end[t]) {
    const uint4 c = d.conn[e];
    // F in FP64 from the FP64 displacement and rest inverse; det F from the
    // displacement form, so inversion is decided in FP64 in both paths
    double u64[12], dm64[9], H64[9];
    gather_vec64(d.u_eval, c, u64);
#pragma unroll
    for (int k = 0; k < 9; ++k) dm64[k] = d.dminv[(size_t)k * d.T + e];
    displacement_gradient(u64, dm64, H64);
    const bool upright = 1.0 + volume_change(H64) > 0.0;
    Real dm[9], g[12], H[9], F[9];
    const uint8_t flags = d.flags[e];
    // a plastic neo-Hookean tet turns with its elastic part Fe = F Fp^-1
    // (Fp^-1 = I + ep, FP64): the operator's shear frame
    const bool flowing = (flags & FLAG_SNH) && (flags & FLAG_PLASTIC);
    double Frot64[9];
    if (flowing) {
      double F64[9], Fi[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        F64[k] = H64[k] + ((k % 4 == 0) ? 1.0 : 0.0);
        Fi[k] = d.plastic[(size_t)k * d.T + e] + ((k % 4 == 0) ? 1.0 : 0.0);
      }
      mul3(F64, Fi, Frot64);
    } else {
#pragma unroll
      for (int k = 0; k < 9; ++k) Frot64[k] = H64[k] + ((k % 4 == 0) ? 1.0 : 0.0);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      dm[k] = Real(dm64[k]);
      H[k] = Real(H64[k]);
      F[k] = Real(Frot64[k]);
    }
    shape_gradients(dm, g);
    // committed rotation: polar in the element precision, sign-fixed SVD in
    // FP64 when inverted; the mixed path re-orthogonalises it in FP64
    Real Rm[9];
    double R64[9];
    if (upright) {
      polar(F, Rm);
#pragma unroll
      for (int i = 0; i < 9; ++i) R64[i] = double(Rm[i]);
      if (MIXED) {
        orthonormalize(R64);
#pragma unroll
        for (int i = 0; i < 9; ++i) Rm[i] = Real(R64[i]);
      }
    } else {
      svd_rotation(Frot64, R64);
#pragma unroll
      for (int i = 0; i < 9; ++i) Rm[i] = Real(R64[i]);
    }
    const Real V = d.vol[e], lam = d.lam[e], mu = d.mu[e];
    Real f[12];
    Real psi;
    Real Wm[9];
    Real lin[8] = {Real(1), Real(0), Real(0), Real(0), Real(0), Real(0), Real(0), Real(0)};
    if (flags & FLAG_SNH) {
      const bool plastic = (flags & FLAG_PLASTIC) != 0;
      Real ep[9], He[9], Fe[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) ep[k] = plastic ? Real(d.plastic[(size_t)k * d.T + e]) : Real(0);
      if (MIXED) {
        // frame indifference, P(Fe) = R P(R^T Fe): the rotation-free
        // R^T Fe - I is formed in FP64 and the law runs in FP32 on it, so a
        // large rotation costs no FP32 digits of the strain
        double G64[9];
        mul3_tn(R64, Frot64, G64);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          He[k] = Real(G64[k] - ((k % 4 == 0) ? 1.0 : 0.0));
          Fe[k] = Real(Frot64[k]);
        }
      } else {
        // H-form at Fe = F Fp^-1: He = H + ep + H ep (Fp^-1 = I + ep)
        Real HE[9];
        mul3(H, ep, HE);
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          He[k] = H[k] + ep[k] + HE[k];
          Fe[k] = He[k] + ((k % 4 == 0) ? Real(1) : Real(0));
        }
      }
      Real Pk[9], PF[9];
      psi = snh_stress(He, lam, mu, Pk);
      if (MIXED) {
        Real RP[9];
        mul3(Rm, Pk, RP);
#pragma unroll
        for (int k = 0; k < 9; ++k) Pk[k] = RP[k];
      }
      // P = P(Fe) Fp^-T and the volume frame cof(Fe) Fp^-T
      Real C[9];
      cof3(Fe, C);
      if (plastic) {
        ep[0] += Real(1);
        ep[4] += Real(1);
        ep[8] += Real(1);
        mul3_nt(Pk, ep, PF);
        mul3_nt(C, ep, Wm);
      } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          PF[k] = Pk[k];
          Wm[k] = C[k];
        }
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) d.W[(size_t)k * d.T + e] = Wm[k];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        Real tmp[3];
        mv3(PF, &g[3 * i], tmp);
        f[3 * i + 0] = -V * tmp[0];
        f[3 * i + 1] = -V * tmp[1];
        f[3 * i + 2] = -V * tmp[2];
      }
      psi *= V;
    } else {
      // corotated: D' = R^T F - I, eps = sym(D'); FP64 in the mixed path
      // (orthogonal R, rotation-free result), else (R^T - I) + R^T H
      Real Dp[9], eps[9], S[9];
      if (MIXED) {
        double F64[9], G64[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F64[k] = H64[k] + ((k % 4 == 0) ? 1.0 : 0.0);
        mul3_tn(R64, F64, G64);
#pragma unroll
        for (int k = 0; k < 9; ++k) Dp[k] = Real(G64[k] - ((k % 4 == 0) ? 1.0 : 0.0));
      } else {
        Real RH[9];
        mul3_tn(Rm, H, RH);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            Dp[r * 3 + k] = (Rm[k * 3 + r] - (r == k ? Real(1) : Real(0))) + RH[r * 3 + k];
      }
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) eps[r * 3 + k] = Real(0.5) * (Dp[r * 3 + k] + Dp[k * 3 + r]);
      double pl[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) pl[k] = d.plastic[(size_t)k * d.T + e];
      if (flags & FLAG_PLASTIC) {
        // tetPointTrial (material.cc:1004-1022): the return map from the
        // committed point, loading past the margin; its linearisation is the
        // return map's when loading, else the elastic law
        Real excess;
        const bool returned = j2_return_lin(eps, pl, lam, mu, Real(d.yield[e]), Real(d.hard[e]),
                                            P.elastic_trial == 0u, S, (double *)nullptr, lin, &excess);
        const bool loading = returned && double(excess) > P.margin;
        if (loading) atomicAdd(&d.loading[b], 1u);
        if (!loading) {
          lin[0] = Real(1);
#pragma unroll
          for (int k = 1; k < 8; ++k) lin[k] = Real(0);
        }
        if (P.newton) {
#pragma unroll
          for (int k = 0; k < 8; ++k) d.lin[(size_t)k * d.T + e] = lin[k];
        }
      } else {
        const Real tr = eps[0] + eps[4] + eps[8];
#pragma unroll
        for (int k = 0; k < 9; ++k) S[k] = Real(2) * mu * eps[k] + ((k % 4 == 0) ? lam * tr : Real(0));
      }
      Real el = Real(0);
#pragma unroll
      for (int k = 0; k < 9; ++k) el += S[k] * (eps[k] - Real(pl[k]));
      psi = Real(0.5) * V * el;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        Real tmp[3], w[3];
        mv3(S, &g[3 * i], tmp);
        mv3(Rm, tmp, w);
        f[3 * i + 0] = -V * w[0];
        f[3 * i + 1] = -V * w[1];
        f[3 * i + 2] = -V * w[2];
      }
    }
    // the damping force needs the rest gradients of the Real law
    const bool damped = DAMP && (flags & FLAG_DAMP);
    Real fit[DAMP ? FIT_PLANES : 1];
    if (damped) damp_assembly(d, P, e, c, u64, g, Rm, V, lam, mu, viscous, f, fit);
    store_slots(d.slot_f, d.slot_pos[e], f);
#pragma unroll
    for (int k = 0; k < 9; ++k) d.R[(size_t)k * d.T + e] = Rm[k];
    if (DIAG) {
      // stiffnessDiagonal: R K_ii R^T = V((lambda+mu)(Rg)(Rg)^T + mu |g|^2 I);
      // neo-Hookean: V(mu (Rg)(Rg)^T + lambda (Wg)(Wg)^T + mu |g|^2 I);
      // a yielding corotated tet under the consistent tangent (material.cc:
      // 1033-1054): V((K + mu theta/3)(Rg)(Rg)^T + mu theta |g|^2 I
      // - 2 mu thetaBar (R n g)(R n g)^T), K = lambda + 2 mu / 3, with the
      // (R n) frame in place of W
      const bool snh = (flags & FLAG_SNH) != 0;
      const bool tangent = P.tangent && !snh && (flags & FLAG_PLASTIC);
      Real lr = V * (snh ? mu : lam + mu), lw = snh ? V * lam : Real(0), vm = V * mu;
      if (tangent) {
        const Real n[9] = {lin[2], lin[3], lin[4], lin[3], lin[5], lin[6], lin[4], lin[6], lin[7]};
        lr = V * (lam + Real(2) * mu / Real(3) + mu * lin[0] / Real(3));
        lw = Real(-2) * V * mu * lin[1];
        vm = V * mu * lin[0];
        mul3(Rm, n, Wm);  // (R n) g
      } else if (!snh) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Wm[k] = Real(0);
      }
      // the shear part along the plastic reference's gradients Fp^-T g
      Real gs[12];
      plastic_gradients(d, e, flowing, g, gs);
      const uint4 pos = d.slot_pos[e];
      const uint32_t at[4] = {pos.x, pos.y, pos.z, pos.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        Real *o = d.slot_d + 6 * (size_t)at[i];
        Real rg[3], wg[3];
        mv3(Rm, &gs[3 * i], rg);
        mv3(Wm, &g[3 * i], wg);
        const Real gg = gs[3 * i] * gs[3 * i] + gs[3 * i + 1] * gs[3 * i + 1] + gs[3 * i + 2] * gs[3 * i + 2];
        o[0] = lr * rg[0] * rg[0] + lw * wg[0] * wg[0] + vm * gg;
        o[1] = lr * rg[0] * rg[1] + lw * wg[0] * wg[1];
        o[2] = lr * rg[0] * rg[2] + lw * wg[0] * wg[2];
        o[3] = lr * rg[1] * rg[1] + lw * wg[1] * wg[1] + vm * gg;
        o[4] = lr * rg[1] * rg[2] + lw * wg[1] * wg[2];
        o[5] = lr * rg[2] * rg[2] + lw * wg[2] * wg[2] + vm * gg;
      }
      if (damped) damp_diagonal(d, P, e, fit, g, Rm, V, lam, mu);
    }
    if (ENERGY) {
      d.energy[e] = psi;
      psi_total = double(psi);
    }
  }
  if (ENERGY) {
    psi_total = block_sum(psi_total, lds);
    if (threadIdx.x == 0) d.epartial[t] = psi_total;
  }
}

// predictorInverts: det F <= 0 at the alpha-point predictor (FP64)
template <class Real>
__global__ __launch_bounds__(TILE) void k_predictor_inverts(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.tt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  if (e >= d.tt_end[t]) return;
  const uint4 c = d.conn[e];
  const uint32_t nn[4] = {c.x, c.y, c.z, c.w};
  double u[12], dm[9], H[9];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const size_t k = 3 * (size_t)nn[i] + r;
      const double x = d.x[k];
      const double shifted = x + P.inner * d.v[k] + P.carry * d.a[k];
      u[3 * i + r] = (1.0 - P.af) * shifted + P.af * x - d.X[k];
    }
#pragma unroll
  for (int k = 0; k < 9; ++k) dm[k] = d.dminv[(size_t)k * d.T + e];
  displacement_gradient(u, dm, H);
  if (!(1.0 + volume_change(H) > 0.0)) atomicOr(&d.inverts[b], 1);
}

// commitPlastic / flowLattice at final positions, FP64
template <class Real>
__global__ __launch_bounds__(TILE) void k_commit(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.tt_body[t];
  if (!d.params[b].active) return;
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  if (e >= d.tt_end[t]) return;
  const uint8_t flags = d.flags[e];
  if (!(flags & FLAG_PLASTIC)) return;
  const uint4 c = d.conn[e];
  double F[9];
  deformation64(d, d.x, e, c, F);
  double pl[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) pl[k] = d.plastic[(size_t)k * d.T + e];
  bool grew = false;
  if (flags & FLAG_SNH) {
    grew = hencky_flow(F, d.mu64[e], d.yield[e], d.hard[e], d.pmax[e], pl, &pl[9]);
  } else {
    double Rm[9];
    if (det3(F) > 0.0)
      polar(F, Rm);
    else
      svd_rotation(F, Rm);
    double Dp[9], eps[9], S[9], next[10];
    mul3_tn(Rm, F, Dp);
    Dp[0] -= 1.0;
    Dp[4] -= 1.0;
    Dp[8] -= 1.0;
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) eps[r * 3 + k] = 0.5 * (Dp[r * 3 + k] + Dp[k * 3 + r]);
    j2_return(eps, pl, d.lam64[e], d.mu64[e], d.yield[e], d.hard[e], S, next);
    grew = next[9] > pl[9];
    for (int k = 0; k < 10; ++k) pl[k] = next[k];
  }
#pragma unroll
  for (int k = 0; k < 10; ++k) d.plastic[(size_t)k * d.T + e] = pl[k];
  if (grew) {
    d.yflag[e] = 1;  // cleared at the start of every step
    atomicAdd(&d.yields[b], 1u);
  }
}

// overstretched: sqrt(lambda_max(F^T F)) > 1 + tearStrain, FP64
template <class Real>
__global__ __launch_bounds__(TILE) void k_tear(Dev<Real> d) {
  const uint32_t t = blockIdx.x, b = d.tt_body[t];
  if (!d.params[b].active) return;
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  if (e >= d.tt_end[t]) return;
  uint8_t torn = 0;
  if (d.flags[e] & FLAG_TEAR) {
    double F[9];
    deformation64(d, d.x, e, d.conn[e], F);
    torn = stretch_of(F, nullptr) > d.tear[e] ? 1 : 0;
  }
  d.tflag[e] = torn;
}

// ------------------------------------------------ kernel-level entry points

template <class Real>
__global__ __launch_bounds__(TILE) void k_displacement(Dev<Real> d, const double *xin) {
  const uint32_t n = blockIdx.x * TILE + threadIdx.x;
  if (n >= d.N) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = 3 * (size_t)n + c;
    d.u_eval[i] = xin[i] - d.X[i];
  }
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_gather_forces(Dev<Real> d, Real *out) {
  const uint32_t n = blockIdx.x * TILE + threadIdx.x;
  if (n >= d.N) return;
  Real f[3] = {Real(0), Real(0), Real(0)};
  const uint32_t k1 = d.inc_start[n + 1];
  for (uint32_t k = d.inc_start[n]; k < k1; ++k) {
    f[0] += d.slot_f[3 * (size_t)k + 0];
    f[1] += d.slot_f[3 * (size_t)k + 1];
    f[2] += d.slot_f[3 * (size_t)k + 2];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[3 * (size_t)n + c] = f[c];
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_apply_plain(Dev<Real> d, const Real *dir, Real *out,
                                                      double inertia, double reach) {
  const uint32_t n = blockIdx.x * TILE + threadIdx.x;
  if (n >= d.N) return;
  Real q[3] = {Real(0), Real(0), Real(0)};
  if (!d.anchored[n]) {
    const uint32_t k1 = d.inc_start[n + 1];
    for (uint32_t k = d.inc_start[n]; k < k1; ++k) {
      q[0] += d.slot_f[3 * (size_t)k + 0];
      q[1] += d.slot_f[3 * (size_t)k + 1];
      q[2] += d.slot_f[3 * (size_t)k + 2];
    }
    const Real im = Real(inertia * d.mass[n]), rc = Real(reach);
#pragma unroll
    for (int c = 0; c < 3; ++c) q[c] = im * dir[3 * (size_t)n + c] + rc * q[c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) out[3 * (size_t)n + c] = q[c];
}

// The operator of one body with its own input, output and slot buffer (the
// allocation-free entry point): body-local FP64 vectors d and y (3 per node
// from the body's first node n0), the body's 4 T_b corner slots at their CSR
// positions less the body's first position k0.  The same operations in the
// same order as k_op_element_flat + k_apply_plain.
template <class Real>
__global__ __launch_bounds__(TILE) void k_apply_body_element(Dev<Real> d, uint32_t t0, uint32_t t1,
                                                             uint32_t n0, uint32_t k0,
                                                             const double *din, Real *slots) {
  const uint32_t e = t0 + blockIdx.x * TILE + threadIdx.x;
  if (e >= t1) return;
  const uint4 c = d.conn[e];
  const uint32_t nn[4] = {c.x - n0, c.y - n0, c.z - n0, c.w - n0};
  Real dv[12], y[12];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) dv[3 * i + r] = Real(din[3 * (size_t)nn[i] + r]);
  op_tet_values(d, e, dv, y);
  const uint4 pos = d.slot_pos[e];
  store_slots(slots, make_uint4(pos.x - k0, pos.y - k0, pos.z - k0, pos.w - k0), y);
}

// The damped operator of one body (a test entry point): each tet writes
// reach K d + damp tau C d to its slots, gathered with reach 1.
template <class Real>
__global__ __launch_bounds__(TILE) void k_apply_body_damped(Dev<Real> d, uint32_t t0, uint32_t t1,
                                                            uint32_t n0, uint32_t k0, const double *din,
                                                            double reach, double damp, Real *slots) {
  const uint32_t e = t0 + blockIdx.x * TILE + threadIdx.x;
  if (e >= t1) return;
  const uint4 c = d.conn[e];
  const uint32_t nn[4] = {c.x - n0, c.y - n0, c.z - n0, c.w - n0};
  Real dv[12], y[12];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) dv[3 * i + r] = Real(din[3 * (size_t)nn[i] + r]);
  op_tet_values(d, e, dv, y);
#pragma unroll
  for (int k = 0; k < 12; ++k) y[k] *= Real(reach);
  if (d.fit && (d.flags[e] & FLAG_DAMP)) op_damping(d, e, dv, Real(damp), y);
  const uint4 pos = d.slot_pos[e];
  store_slots(slots, make_uint4(pos.x - k0, pos.y - k0, pos.z - k0, pos.w - k0), y);
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_apply_body_gather(Dev<Real> d, uint32_t n0, uint32_t n1,
                                                            uint32_t k0, const Real *slots,
                                                            const double *din, double *yout,
                                                            double inertia, double reach) {
  const uint32_t n = n0 + blockIdx.x * TILE + threadIdx.x;
  if (n >= n1) return;
  const size_t l = n - n0;
  Real q[3] = {Real(0), Real(0), Real(0)};
  if (!d.anchored[n]) {
    const uint32_t k1 = d.inc_start[n + 1] - k0;
    for (uint32_t k = d.inc_start[n] - k0; k < k1; ++k) {
      q[0] += slots[3 * (size_t)k + 0];
      q[1] += slots[3 * (size_t)k + 1];
      q[2] += slots[3 * (size_t)k + 2];
    }
    const Real im = Real(inertia * d.mass[n]), rc = Real(reach);
#pragma unroll
    for (int c = 0; c < 3; ++c) q[c] = im * Real(din[3 * l + c]) + rc * q[c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) yout[3 * l + c] = double(q[c]);
}

// The operator with the initial-stress term (the static Newton's
// scratch.stressed, solver.cc:947-948) of one body at positions xin (FP64,
// body-local), after k_assemble at those positions: each tet reads its
// corner forces from its slots, adds initialStress d (FP64) to the
// operator's R K R^T d for corotated tets, and writes the result back to
// the same slots.
template <class Real>
__global__ __launch_bounds__(TILE) void k_stressed_element(Dev<Real> d, uint32_t t0, uint32_t t1, uint32_t n0,
                                                           uint32_t k0, const double *xin, const double *din,
                                                           Real *slots) {
  const uint32_t e = t0 + blockIdx.x * TILE + threadIdx.x;
  if (e >= t1) return;
  const uint4 c = d.conn[e];
  const uint32_t nn[4] = {c.x - n0, c.y - n0, c.z - n0, c.w - n0};
  const uint32_t gn[4] = {c.x, c.y, c.z, c.w};
  Real dv[12], y[12];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) dv[3 * i + r] = Real(din[3 * (size_t)nn[i] + r]);
  op_tet_values(d, e, dv, y);
  const uint4 pos = d.slot_pos[e];
  const uint4 at = make_uint4(pos.x - k0, pos.y - k0, pos.z - k0, pos.w - k0);
  if (!(d.flags[e] & FLAG_SNH)) {
    const uint32_t ps[4] = {at.x, at.y, at.z, at.w};
    double x[12], X[12], f[12], dd[12], dm[9], g[12], R[9], tw[36], tu[36], add[12];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        x[3 * i + r] = xin[3 * (size_t)nn[i] + r];
        X[3 * i + r] = d.X[3 * (size_t)gn[i] + r];
        f[3 * i + r] = double(slots[3 * (size_t)ps[i] + r]);
        dd[3 * i + r] = din[3 * (size_t)nn[i] + r];
        add[3 * i + r] = 0.0;
      }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      dm[k] = d.dminv[(size_t)k * d.T + e];
      R[k] = double(d.R[(size_t)k * d.T + e]);
    }
    shape_gradients(dm, g);
    if (initial_stress_tet(x, X, g, d.vol64[e], d.lam64[e], d.mu64[e], R, f, tw, tu)) {
      apply_initial_stress(tw, tu, dd, add);
#pragma unroll
      for (int k = 0; k < 12; ++k) y[k] += Real(add[k]);
    }
  }
  store_slots(slots, at, y);
}

inline unsigned blocks_for(uint32_t n) { return (n + TILE - 1) / TILE; }

// ------------------------------------------------- corrections converge

// per node tile: max |dx| and max |a| over the tile's free nodes
// Plastic Newton, per node tile of an active body: sum of |right|^2 over
// its nodes (anchored rows are zero), the largest node increment |dx| and
// the largest node position |x_eval| over its free nodes (the reference's
// glm::length per node).
template <class Real>
__global__ __launch_bounds__(TILE) void k_newton_norms(Dev<Real> d) {
  __shared__ double lds[3][TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double sq = 0.0, moved = 0.0, extent = 0.0;
  if (n < d.nt_end[t]) {
    double dd = 0.0, pp = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t i = 3 * (size_t)n + c;
      const double r = d.right[i], dx = double(d.dx[i]), p = d.u_eval[i] + d.X[i];
      sq += r * r;
      dd += dx * dx;
      pp += p * p;
    }
    if (!d.anchored[n]) {
      moved = sqrt(dd);
      extent = sqrt(pp);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sq += __shfl_down(sq, off, 64);
    moved = fmax(moved, __shfl_down(moved, off, 64));
    extent = fmax(extent, __shfl_down(extent, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    lds[0][wave] = sq;
    lds[1][wave] = moved;
    lds[2][wave] = extent;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (int w = 0; w < TILE / 64; ++w) {
      s0 += lds[0][w];
      m1 = fmax(m1, lds[1][w]);
      m2 = fmax(m2, lds[2][w]);
    }
    d.nnorm[3 * t] = s0;
    d.nnorm[3 * t + 1] = m1;
    d.nnorm[3 * t + 2] = m2;
  }
}

// bodyRotation (factor.cc:615-646).  The reference positions of the bodies
// whose factor was rebuilt are their positions now.
template <class Real>
__global__ __launch_bounds__(TILE) void k_frame_reference(Dev<Real> d, const int *refresh) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!refresh[b]) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) d.fref[3 * (size_t)n + c] = d.x[3 * (size_t)n + c];
}

// Per node tile of an active body: mass, mass-weighted positions now and at
// the factorization (7 sums; mode 0), or the mass-weighted correlation of
// the centred positions, sum m (x - c)(r - c_r)^T (9 sums; mode 1).
template <class Real>
__global__ __launch_bounds__(TILE) void k_frame_moments(Dev<Real> d, int mode) {
  __shared__ double lds[TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (n < d.nt_end[t]) {
    const double m = d.mass[n];
    const double *x = d.x + 3 * (size_t)n, *r = d.fref + 3 * (size_t)n;
    if (mode == 0) {
      v[0] = m;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        v[1 + c] = m * x[c];
        v[4 + c] = m * r[c];
      }
    } else {
      const double *cen = d.fcent + 6 * (size_t)b;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) v[3 * i + j] = m * (x[i] - cen[i]) * (r[j] - cen[3 + j]);
    }
  }
  const int count = mode == 0 ? 7 : 9;
  for (int k = 0; k < count; ++k) {
    const double total = block_sum(v[k], lds);
    if (threadIdx.x == 0) d.fpart[9 * (size_t)t + k] = total;
    __syncthreads();
  }
}

// One block per active body: its tiles' sums in tile order, then the
// centroids (mode 0) or the frame, polarRotation of the correlation
// (mode 1; identity when the correlation is singular, as the reference).
template <class Real>
__global__ __launch_bounds__(TILE) void k_frame_body(Dev<Real> d, int mode) {
  __shared__ double lds[TILE / 64];
  const uint32_t b = blockIdx.x;
  if (!d.params[b].active) return;
  const uint32_t t0 = d.body_nt0[b], t1 = d.body_nt1[b];
  double sum[9];
  const int count = mode == 0 ? 7 : 9;
  for (int k = 0; k < count; ++k) {
    double s = 0.0;
    for (uint32_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) s += d.fpart[9 * (size_t)t + k];
    sum[k] = block_sum(s, lds);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    const double m = sum[0] > 0.0 ? sum[0] : 1.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) d.fcent[6 * (size_t)b + k] = sum[1 + k] / m;
  } else {
    double R[9];
    polar(sum, R);
#pragma unroll
    for (int k = 0; k < 9; ++k) d.frame[9 * (size_t)b + k] = R[k];
  }
}

// F^T r into rf for the nodes of bodies still iterating
template <class Real>
__global__ __launch_bounds__(TILE) void k_frame_in(Dev<Real> d, int parity) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active || (parity >= 0 && d.done[parity * d.B + b])) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
  const double *F = d.frame + 9 * (size_t)b;
  const Real *r = d.r + 3 * (size_t)n;
  const double r0 = double(r[0]), r1 = double(r[1]), r2 = double(r[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c) d.rf[3 * (size_t)n + c] = Real(F[c] * r0 + F[3 + c] * r1 + F[6 + c] * r2);
}

// Iterate updates of the plastic Newton on active bodies (NewtonVec).
template <class Real>
__global__ __launch_bounds__(TILE) void k_newton_vec(Dev<Real> d, int mode) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
  const bool free = !d.anchored[n];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = 3 * (size_t)n + c;
    switch (mode) {
      case NV_BACKTRACK: {
        const Real half = d.dx[i] * Real(0.5);
        d.dx[i] = half;
        d.a[i] -= double(half);
        break;
      }
      case NV_SAVE: d.lowest[i] = d.a[i]; break;
      case NV_RESTORE: d.a[i] = d.lowest[i]; break;
      case NV_RESET:
        d.a[i] = d.carried[i];
        d.x0[i] = Real(d.carried[i]);
        d.u_eval[i] = d.weighted[i] - d.X[i];
        break;
      case NV_DX_FROM_A: d.dx[i] = free ? Real(d.a[i]) : Real(0); break;
      default: break;
    }
  }
}

template <class Real>
__global__ __launch_bounds__(TILE) void k_increment_norms(Dev<Real> d, double *out) {
  __shared__ double lds[2][TILE / 64];
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  double dx = 0.0, a = 0.0;
  if (d.params[b].active && n < d.nt_end[t] && !d.anchored[n]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      dx = fmax(dx, fabs(double(d.dx[3 * (size_t)n + c])));
      a = fmax(a, fabs(d.a[3 * (size_t)n + c]));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dx = fmax(dx, __shfl_down(dx, off, 64));
    a = fmax(a, __shfl_down(a, off, 64));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    lds[0][wave] = dx;
    lds[1][wave] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m0 = 0.0, m1 = 0.0;
    for (int w = 0; w < TILE / 64; ++w) {
      m0 = fmax(m0, lds[0][w]);
      m1 = fmax(m1, lds[1][w]);
    }
    out[2 * t] = m0;
    out[2 * t + 1] = m1;
  }
}

// ------------------------------------------------------------- adjoint

// d cof(F)[D]: the columns of cof F are f1 x f2, f2 x f0, f0 x f1 (0-based
// columns of F), so each column's derivative is two cross products.
__device__ __forceinline__ void cof_dir(const double *F, const double *D, double *dC) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int a = (c + 1) % 3, q = (c + 2) % 3;
    const double fa[3] = {F[a], F[3 + a], F[6 + a]}, fb[3] = {F[q], F[3 + q], F[6 + q]};
    const double da[3] = {D[a], D[3 + a], D[6 + a]}, db[3] = {D[q], D[3 + q], D[6 + q]};
    dC[0 * 3 + c] = (da[1] * fb[2] - da[2] * fb[1]) + (fa[1] * db[2] - fa[2] * db[1]);
    dC[1 * 3 + c] = (da[2] * fb[0] - da[0] * fb[2]) + (fa[2] * db[0] - fa[0] * db[2]);
    dC[2 * 3 + c] = (da[0] * fb[1] - da[1] * fb[0]) + (fa[0] * db[1] - fa[1] * db[0]);
  }
}

// SNH stress derivative dP/dF[D] = mu (D - dcof) + (lambda+mu)((C:D) C + (J-1) dcof)
__device__ __forceinline__ void snh_dP(const double *F, const double *D, double lam, double mu,
                                       double *dP) {
  double C[9], dC[9];
  cof3(F, C);
  cof_dir(F, D, dC);
  const double J = det3(F);
  double cd = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) cd += C[i] * D[i];
#pragma unroll
  for (int i = 0; i < 9; ++i) dP[i] = mu * (D[i] - dC[i]) + (lam + mu) * (cd * C[i] + (J - 1.0) * dC[i]);
}

// F of tet e at the evaluation point (u_eval) and its FP64 shape gradients
template <class Real>
__device__ __forceinline__ void eval_frame(const Dev<Real> &d, uint32_t e, uint4 c, double *F,
                                           double *dm, double *g) {
  double u[12];
  gather_vec64(d.u_eval, c, u);
#pragma unroll
  for (int k = 0; k < 9; ++k) dm[k] = d.dminv[(size_t)k * d.T + e];
  displacement_gradient(u, dm, F);
  F[0] += 1.0;
  F[4] += 1.0;
  F[8] += 1.0;
  shape_gradients(dm, g);
}

// y_i = V dP(D) g_i, D = sum_j dir_j g_j^T: the neo-Hookean Hessian at the
// evaluation point applied to dir, per tet into the force slots
template <class Real>
__global__ __launch_bounds__(TILE) void k_hvp(Dev<Real> d, const Real *dir, int parity) {
  const uint32_t t = blockIdx.x, b = d.tt_body[t];
  if (!d.params[b].active || (parity >= 0 && d.done[parity * d.B + b])) return;
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  if (e >= d.tt_end[t]) return;
  const uint4 c = d.conn[e];
  double F[9], dm[9], g[12], dv[12], D[9], dP[9];
  eval_frame(d, e, c, F, dm, g);
  const uint32_t n[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) dv[3 * i + r] = double(dir[3 * (size_t)n[i] + r]);
  displacement_gradient(dv, dm, D);
  snh_dP(F, D, d.lam64[e], d.mu64[e], dP);
  const double V = d.vol64[e];
  Real y[12];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double tmp[3];
    mv3(dP, &g[3 * i], tmp);
#pragma unroll
    for (int r = 0; r < 3; ++r) y[3 * i + r] = Real(V * tmp[r]);
  }
  store_slots(d.slot_f, d.slot_pos[e], y);
}

// x_w = x_n + (1-af)(h v_n + c a_n) + reach a' (as u_eval = x_w - X), the
// adjoint right-hand side abar + beta h^2 xbar + gamma h vbar on free rows,
// zero initial guess
template <class Real>
__global__ __launch_bounds__(TILE) void k_adj_prepare(Dev<Real> d, AdjDev a) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = 3 * (size_t)n + c;
    const double xw = a.xn[i] + (1.0 - P.af) * (P.inner * a.vn[i] + P.carry * a.an[i]) + P.reach * a.anext[i];
    d.u_eval[i] = xw - d.X[i];
    d.right[i] = d.anchored[n] ? 0.0
                               : a.ab[i] + P.beta * P.inner * P.inner * a.xb[i] + P.gamma * P.inner * a.vb[i];
    d.x0[i] = Real(0);
  }
}

// K lambda over every row (slots of k_hvp on lambda = dx, fixed CSR order),
// the mass-term weight, and the adjoint state one substep back
template <class Real>
__global__ __launch_bounds__(TILE) void k_adj_update(Dev<Real> d, AdjDev a) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  const StepParams &P = d.params[b];
  if (!P.active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
  double kl[3] = {0.0, 0.0, 0.0};
  const uint32_t k1 = d.inc_start[n + 1];
  for (uint32_t k = d.inc_start[n]; k < k1; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) kl[c] += double(d.slot_f[3 * (size_t)k + c]);
  }
  const bool fixed = d.anchored[n] != 0;
  const double m = d.mass[n];
  double w = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t i = 3 * (size_t)n + c;
    const double lam = fixed ? 0.0 : double(d.dx[i]);
    const double x1 = a.xb[i], v1 = a.vb[i];
    a.kl[i] = kl[c];
    w += (P.inertia * a.anext[i] + P.am * a.an[i] - P.g[c]) * lam;
    a.xb[i] = x1 - kl[c];
    if (fixed) {
      a.vb[i] = -(1.0 - P.af) * P.inner * kl[c];
      a.ab[i] = -(1.0 - P.af) * P.carry * kl[c];
    } else {
      a.vb[i] = P.inner * x1 + v1 - (1.0 - P.af) * P.inner * kl[c];
      a.ab[i] = P.carry * x1 + P.inner * (1.0 - P.gamma) * v1 - P.am * m * lam -
                (1.0 - P.af) * P.carry * kl[c];
    }
  }
  a.w[n] = fixed ? 0.0 : w;
}

// per tet at x_w with D = sum_j lambda_j g_j^T:
//   lambda_bar -= V (J-1) C:D,  mu_bar -= V ((F - C) + (J-1) C):D,
//   rest slots: -dPhi/dX with dPhi/dDm = V[(P:D) I - F^T dP(D) - D^T P] Dm^-T
//   and the mass term (rho/4) sum_{i in e} w_i V Dm^-T
template <class Real>
__global__ __launch_bounds__(TILE) void k_adj_params(Dev<Real> d, AdjDev a) {
  const uint32_t t = blockIdx.x, b = d.tt_body[t];
  if (!d.params[b].active) return;
  const uint32_t e = d.tt_begin[t] + threadIdx.x;
  if (e >= d.tt_end[t]) return;
  const uint4 c = d.conn[e];
  const uint32_t n[4] = {c.x, c.y, c.z, c.w};
  double F[9], dm[9], g[12], le[12], D[9], C[9], P[9], dP[9];
  eval_frame(d, e, c, F, dm, g);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 3; ++r) le[3 * i + r] = d.anchored[n[i]] ? 0.0 : double(d.dx[3 * (size_t)n[i] + r]);
  displacement_gradient(le, dm, D);
  cof3(F, C);
  const double J = det3(F), lam = d.lam64[e], mu = d.mu64[e], V = d.vol64[e];
  double pl = 0.0, pm = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    pl += (J - 1.0) * C[i] * D[i];
    pm += ((F[i] - C[i]) + (J - 1.0) * C[i]) * D[i];
    P[i] = mu * (F[i] - C[i]) + (lam + mu) * (J - 1.0) * C[i];
  }
  a.lam_bar[e] -= V * pl;
  a.mu_bar[e] -= V * pm;
  snh_dP(F, D, lam, mu, dP);
  double PD = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) PD += P[i] * D[i];
  const double we = 0.25 * d.rho[e] * (a.w[n[0]] + a.w[n[1]] + a.w[n[2]] + a.w[n[3]]);
  double A[9], G[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double fq = 0.0, dp = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        fq += F[k * 3 + r] * dP[k * 3 + q];
        dp += D[k * 3 + r] * P[k * 3 + q];
      }
      A[r * 3 + q] = V * ((r == q ? PD + we : 0.0) - fq - dp);
    }
  mul3_nt(A, dm, G);
  const uint4 pos = d.slot_pos[e];
  double *o0 = a.xslot + 3 * (size_t)pos.x, *o1 = a.xslot + 3 * (size_t)pos.y,
         *o2 = a.xslot + 3 * (size_t)pos.z, *o3 = a.xslot + 3 * (size_t)pos.w;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    o0[r] = G[r * 3 + 0] + G[r * 3 + 1] + G[r * 3 + 2];
    o1[r] = -G[r * 3 + 0];
    o2[r] = -G[r * 3 + 1];
    o3[r] = -G[r * 3 + 2];
  }
}

// X_bar += per-node sums of the rest slots in CSR order
template <class Real>
__global__ __launch_bounds__(TILE) void k_adj_gather_rest(Dev<Real> d, AdjDev a) {
  const uint32_t t = blockIdx.x, b = d.nt_body[t];
  if (!d.params[b].active) return;
  const uint32_t n = d.nt_begin[t] + threadIdx.x;
  if (n >= d.nt_end[t]) return;
  double s3[3] = {0.0, 0.0, 0.0};
  const uint32_t k1 = d.inc_start[n + 1];
  for (uint32_t k = d.inc_start[n]; k < k1; ++k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) s3[c] += a.xslot[3 * (size_t)k + c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) a.X_bar[3 * (size_t)n + c] += s3[c];
}

}  // namespace

// ---------------------------------------------------------------- launchers

#define TMI_LAUNCH(kernel, grid, s, ...) \
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(TILE), 0, s, __VA_ARGS__)

template <class Real>
void Launch<Real>::predict(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_predict<Real>, d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::predictor_inverts(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_predictor_inverts<Real>, d.n_ttiles, s, d);
}
template <class Real>
void Launch<Real>::evaluate(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_evaluate<Real>, d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::assemble(const Dev<Real> &d, bool diag, bool energy, hipStream_t s, int viscous) {
  if (d.fit) {
    if (diag && energy) TMI_LAUNCH((k_assemble<Real, true, true, true>), d.n_ttiles, s, d, viscous);
    else if (diag) TMI_LAUNCH((k_assemble<Real, true, false, true>), d.n_ttiles, s, d, viscous);
    else if (energy) TMI_LAUNCH((k_assemble<Real, false, true, true>), d.n_ttiles, s, d, viscous);
    else TMI_LAUNCH((k_assemble<Real, false, false, true>), d.n_ttiles, s, d, viscous);
    return;
  }
  if (diag && energy) TMI_LAUNCH((k_assemble<Real, true, true, false>), d.n_ttiles, s, d, viscous);
  else if (diag) TMI_LAUNCH((k_assemble<Real, true, false, false>), d.n_ttiles, s, d, viscous);
  else if (energy) TMI_LAUNCH((k_assemble<Real, false, true, false>), d.n_ttiles, s, d, viscous);
  else TMI_LAUNCH((k_assemble<Real, false, false, false>), d.n_ttiles, s, d, viscous);
}
template <class Real>
void Launch<Real>::balance(const Dev<Real> &d, bool diag, bool taken, hipStream_t s) {
  if (diag && taken) TMI_LAUNCH((k_balance<Real, true, true>), d.n_ntiles, s, d);
  else if (diag) TMI_LAUNCH((k_balance<Real, true, false>), d.n_ntiles, s, d);
  else if (taken) TMI_LAUNCH((k_balance<Real, false, true>), d.n_ntiles, s, d);
  else TMI_LAUNCH((k_balance<Real, false, false>), d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::pcg_init(const Dev<Real> &d, bool with_x0, bool factor, hipStream_t s) {
  if (with_x0 && factor) TMI_LAUNCH((k_pcg_init<Real, true, true>), d.n_ntiles, s, d);
  else if (with_x0) TMI_LAUNCH((k_pcg_init<Real, true, false>), d.n_ntiles, s, d);
  else if (factor) TMI_LAUNCH((k_pcg_init<Real, false, true>), d.n_ntiles, s, d);
  else TMI_LAUNCH((k_pcg_init<Real, false, false>), d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::pcg_rz(const Dev<Real> &d, bool start, int parity, hipStream_t s) {
  if (start) TMI_LAUNCH((k_pcg_rz<Real, true>), d.n_ntiles, s, d, parity);
  else TMI_LAUNCH((k_pcg_rz<Real, false>), d.n_ntiles, s, d, parity);
}
template <class Real>
void Launch<Real>::pcg_begin(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_pcg_begin<Real>, d.B, s, d);
}
template <class Real>
void Launch<Real>::op_element(const Dev<Real> &d, const Real *dir, int parity, hipStream_t s) {
  const size_t lds = 3 * size_t(std::max(d.tn_max, 1u)) * sizeof(Real);
  if (d.fit)
    hipLaunchKernelGGL((k_op_element<Real, true>), dim3(d.n_ttiles), dim3(TILE), lds, s, d, dir, parity);
  else
    hipLaunchKernelGGL((k_op_element<Real, false>), dim3(d.n_ttiles), dim3(TILE), lds, s, d, dir, parity);
}
template <class Real>
void Launch<Real>::op_gather(const Dev<Real> &d, const Real *dir, Real *out, int parity,
                             hipStream_t s) {
  TMI_LAUNCH(k_op_gather<Real>, d.n_ntiles, s, d, dir, out, parity);
}
template <class Real>
void Launch<Real>::pcg_update(const Dev<Real> &d, int parity, bool factor, hipStream_t s) {
  if (factor) TMI_LAUNCH((k_pcg_update<Real, true>), d.n_ntiles, s, d, parity);
  else TMI_LAUNCH((k_pcg_update<Real, false>), d.n_ntiles, s, d, parity);
}
template <class Real>
void Launch<Real>::pcg_direction(const Dev<Real> &d, int parity, hipStream_t s) {
  TMI_LAUNCH(k_pcg_direction<Real>, d.n_ntiles, s, d, parity);
}
template <class Real>
void Launch<Real>::apply_increment(const Dev<Real> &d, bool, hipStream_t s) {
  TMI_LAUNCH(k_apply_increment<Real>, d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::track_increment(const Dev<Real> &d, bool add, hipStream_t s) {
  TMI_LAUNCH(k_track_increment<Real>, d.n_ntiles, s, d, add ? 1 : 0);
}
template <class Real>
void Launch<Real>::refine_residual(const Dev<Real> &d, const double *inc, double tolerance, hipStream_t s) {
  TMI_LAUNCH(k_op_element64<Real>, d.n_ttiles, s, d, inc);
  TMI_LAUNCH(k_residual64<Real>, d.n_ntiles, s, d, inc);
  TMI_LAUNCH(k_refine_check<Real>, d.B, s, d, tolerance);
}
template <class Real>
void Launch<Real>::finish(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_finish<Real>, d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::commit(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_commit<Real>, d.n_ttiles, s, d);
}
template <class Real>
void Launch<Real>::tear(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_tear<Real>, d.n_ttiles, s, d);
}
template <class Real>
void Launch<Real>::increment_norms(const Dev<Real> &d, double *out, hipStream_t s) {
  TMI_LAUNCH(k_increment_norms<Real>, d.n_ntiles, s, d, out);
}
template <class Real>
void Launch<Real>::newton_norms(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_newton_norms<Real>, d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::frame_reference(const Dev<Real> &d, const int *refresh, hipStream_t s) {
  TMI_LAUNCH(k_frame_reference<Real>, d.n_ntiles, s, d, refresh);
}
template <class Real>
void Launch<Real>::frames(const Dev<Real> &d, hipStream_t s) {
  for (int mode = 0; mode < 2; ++mode) {
    TMI_LAUNCH(k_frame_moments<Real>, d.n_ntiles, s, d, mode);
    TMI_LAUNCH(k_frame_body<Real>, d.B, s, d, mode);
  }
}
template <class Real>
void Launch<Real>::frame_in(const Dev<Real> &d, int parity, hipStream_t s) {
  TMI_LAUNCH(k_frame_in<Real>, d.n_ntiles, s, d, parity);
}
template <class Real>
void Launch<Real>::newton_vec(const Dev<Real> &d, int mode, hipStream_t s) {
  TMI_LAUNCH(k_newton_vec<Real>, d.n_ntiles, s, d, mode);
}
template <class Real>
void Launch<Real>::hvp(const Dev<Real> &d, const Real *dir, int parity, hipStream_t s) {
  TMI_LAUNCH(k_hvp<Real>, d.n_ttiles, s, d, dir, parity);
}
template <class Real>
void Launch<Real>::adj_prepare(const Dev<Real> &d, const AdjDev &a, hipStream_t s) {
  TMI_LAUNCH(k_adj_prepare<Real>, d.n_ntiles, s, d, a);
}
template <class Real>
void Launch<Real>::adj_update(const Dev<Real> &d, const AdjDev &a, hipStream_t s) {
  TMI_LAUNCH(k_adj_update<Real>, d.n_ntiles, s, d, a);
}
template <class Real>
void Launch<Real>::adj_params(const Dev<Real> &d, const AdjDev &a, hipStream_t s) {
  TMI_LAUNCH(k_adj_params<Real>, d.n_ttiles, s, d, a);
  TMI_LAUNCH(k_adj_gather_rest<Real>, d.n_ntiles, s, d, a);
}
template <class Real>
void Launch<Real>::kinetic(const Dev<Real> &d, hipStream_t s) {
  TMI_LAUNCH(k_kinetic<Real>, d.n_ntiles, s, d);
}
template <class Real>
void Launch<Real>::displacement(const Dev<Real> &d, const double *xin, hipStream_t s) {
  TMI_LAUNCH(k_displacement<Real>, blocks_for(d.N), s, d, xin);
}
template <class Real>
void Launch<Real>::apply_plain(const Dev<Real> &d, const Real *dir, Real *out, double inertia,
                               double reach, hipStream_t s) {
  TMI_LAUNCH(k_apply_plain<Real>, blocks_for(d.N), s, d, dir, out, inertia, reach);
}
template <class Real>
void Launch<Real>::apply_body(const Dev<Real> &d, uint32_t t0, uint32_t t1, uint32_t n0, uint32_t n1,
                              uint32_t k0, const double *din, double *yout, double inertia,
                              double reach, Real *slots, hipStream_t s) {
  if (t1 > t0) TMI_LAUNCH(k_apply_body_element<Real>, blocks_for(t1 - t0), s, d, t0, t1, n0, k0, din, slots);
  if (n1 > n0)
    TMI_LAUNCH(k_apply_body_gather<Real>, blocks_for(n1 - n0), s, d, n0, n1, k0, slots, din, yout, inertia,
               reach);
}

template <class Real>
void Launch<Real>::apply_body_damped(const Dev<Real> &d, uint32_t t0, uint32_t t1, uint32_t n0, uint32_t n1,
                                     uint32_t k0, const double *din, double *yout, double inertia, double reach,
                                     double damp, Real *slots, hipStream_t s) {
  if (t1 > t0)
    TMI_LAUNCH(k_apply_body_damped<Real>, blocks_for(t1 - t0), s, d, t0, t1, n0, k0, din, reach, damp, slots);
  if (n1 > n0)
    TMI_LAUNCH(k_apply_body_gather<Real>, blocks_for(n1 - n0), s, d, n0, n1, k0, slots, din, yout, inertia, 1.0);
}

template <class Real>
void Launch<Real>::apply_body_stressed(const Dev<Real> &d, uint32_t t0, uint32_t t1, uint32_t n0, uint32_t n1,
                                       uint32_t k0, const double *xin, const double *din, double *yout,
                                       double inertia, double reach, Real *slots, hipStream_t s) {
  if (t1 > t0)
    TMI_LAUNCH(k_stressed_element<Real>, blocks_for(t1 - t0), s, d, t0, t1, n0, k0, xin, din, slots);
  if (n1 > n0)
    TMI_LAUNCH(k_apply_body_gather<Real>, blocks_for(n1 - n0), s, d, n0, n1, k0, slots, din, yout, inertia,
               reach);
}

template <class Real>
void Launch<Real>::op_element_plain(const Dev<Real> &d, const Real *dir, hipStream_t s) {
  TMI_LAUNCH(k_op_element_flat<Real>, blocks_for(d.T), s, d, dir);
}
template <class Real>
void Launch<Real>::gather_forces(const Dev<Real> &d, Real *out, hipStream_t s) {
  TMI_LAUNCH(k_gather_forces<Real>, blocks_for(d.N), s, d, out);
}

template struct Launch<float>;
template struct Launch<double>;

}  // namespace tmi
