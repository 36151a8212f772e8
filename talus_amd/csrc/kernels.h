// kernels.h -- device data layout and kernel launchers of libtalus_mi.
//
// Layout in HBM (all bodies concatenated; each body owns a contiguous node
// range and tet range):
//   node vectors      xyz-interleaved, FP64 master state (x, v, a, ...) and
//                     Real working vectors (displacements, PCG vectors)
//   tet connectivity  uint4 per tet (one 16-byte load)
//   Dm^-1             FP64, 9 SoA planes dminv[k*T + e] (coalesced)
//   V, lambda, mu     Real per tet
//   rotations R       Real, 9 SoA planes
//   force slots       Real, 3 per (tet, corner), stored in node order: the
//                     corner c of tet e writes slot_f[3*pos + r] with pos =
//                     slot_pos[e].c, its position in the incidence CSR, so a
//                     node's gather is one contiguous run
//   incidence CSR     node n owns positions inc_start[n]..inc_start[n+1]-1,
//                     one per incident (tet, corner) in ascending 4e+c order,
//                     summed in that fixed order (deterministic, no atomics)
// Work is cut into tiles of TILE items of one body (node tiles and tet
// tiles); block b processes tile b, and per-body scalars (PCG alpha/beta,
// convergence) are reduced from per-tile partial sums in a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tmi {

constexpr int TILE = 256;
constexpr int NPART = 4;  // partial sums per node tile

// Per-body integrator constants for the current substep pass.
struct StepParams {
  double inner;    // substep h
  double carry;    // (1/2 - beta) h^2
  double reach;    // h_eff^2 = (1 - alpha_f) beta h^2
  double inertia;  // 1 - alpha_m
  double am, af, beta, gamma;
  double g[3];
  double tolerance;
  uint32_t active;  // body takes part in this pass
  uint32_t limit;   // CG iteration cap
  // plastic Newton of bodies with yielding corotated lanes (the step's
  // plastic branch): the assembly stores the return maps' linearisations
  // (newton), binds the consistent tangent in the operator and diagonal
  // blocks (tangent), skips the return map (elastic_trial) and counts the
  // tets loading past `margin`
  uint32_t newton, tangent, elastic_trial, pad;
  double margin;
  // stiffness damping (solver.cc:3586-3592, 3746-3783): the operator's
  // damping coefficient (1 - alpha_f) gamma h and the alpha-point lag
  // (1 - alpha_f) h of the viscous velocity v + lag((1 - gamma) a_n + gamma a)
  double damp, lag;
};

// The viscous force subtracted from an assembly's corner forces: none (an
// evaluation of forces alone), at the predictor's velocity (balance without
// an acceleration, solver.cc:3917), or with the current acceleration taken
// (balance(&acceleration), 3931-3985).
enum Viscous : int { VISCOUS_NONE = 0, VISCOUS_PREDICTOR = 1, VISCOUS_TAKEN = 2 };

template <class Real>
struct Dev {
  uint32_t N, T, B;
  // tiles
  const uint32_t *nt_body, *nt_begin, *nt_end;
  uint32_t n_ntiles;
  const uint32_t *tt_body, *tt_begin, *tt_end;
  uint32_t n_ttiles;
  const uint32_t *body_nt0, *body_nt1;
  const StepParams *params;
  // nodes
  const double *X, *mass;
  const uint8_t *anchored;
  const uint32_t *inc_start;
  const uint4 *slot_pos;  // per tet: CSR positions of its 4 corner slots
  double *x, *v, *a, *carried, *shifted, *weighted, *right;
  double *u_eval;  // displacement at the evaluation point (FP64 in both paths)
  Real *P, *x0;
  Real *dx, *r, *z, *p, *q;
  // tets
  const uint4 *conn;
  // per tet tile: its tets' nodes once each, ascending (tn_node[tn_start[t]
  // .. tn_start[t+1]), at most tn_max per tile), and per tet its corners as
  // 16-bit positions in its tile's list (conn_l), so an element kernel can
  // stage the tile's node vectors in LDS and gather from there
  const uint32_t *tn_start, *tn_node;
  const uint2 *conn_l;
  uint32_t tn_max;
  const double *dminv;
  const Real *dminv_r;  // Dm^-1 in the element precision (the operator's copy)
  const Real *vol, *lam, *mu;
  const double *vol64, *lam64, *mu64;
  const double *rho;  // density when masses follow rho V / 4, else 0
  const uint8_t *flags;  // bit0 neo-Hookean, bit1 plastic lane, bit2 tearable
  const double *yield, *hard, *pmax, *tear;
  Real *R;
  Real *W;          // neo-Hookean volume frame cof(Fe) Fp^-T, 9 SoA planes
  double *plastic;  // 10 SoA planes
  // stiffness damping: tau per tet (0 undamped) and the rigid fit of each
  // damped tet at its last assembly, FIT_PLANES SoA planes; both null when
  // no tet is damped
  const Real *tau;
  Real *fit;
  Real *lin;        // J2 linearisation {theta, thetaBar, n (6)}, 8 SoA planes
  uint32_t *loading;  // [B] tets loading past the margin in the last assembly
  double *lowest;   // 3N, the Newton's lowest-residual iterate
  double *nnorm;    // 3 per node tile: sum |right|^2, max |dx|, max |x_eval|
  // bodyRotation: positions at each body's last factorization, the body
  // frame F (9 per body), per node tile moments (9) and per body centroids (6)
  double *fref, *frame, *fpart, *fcent;
  Real *rf;         // F^T r, the preconditioner's input in the factor's frame
  Real *slot_f, *slot_d, *energy;
  uint8_t *yflag, *tflag;
  // reductions
  double *partial;   // NPART per node tile
  double *epartial;  // 1 per tet tile
  // PCG, per body
  double *sq;        // [2][B]
  int *done;         // [2][B]
  uint32_t *iters;   // [B]
  double *target2;   // [B]
  double *rr_last;   // [B]
  int *inverts;      // [B]
  uint32_t *yields;  // [B]
  // FP64 iterative refinement (allocated on first use): corner slots of the
  // FP64 operator (12 per tet), the FP64 residual (3N), the FP64 sum of a
  // correction solve's increments (3N), per node tile |residual|^2 and
  // |right|^2, per body the residual relative to |right| and whether it
  // meets the refinement tolerance
  double *slot64, *rres, *rpart, *rrel, *rbase;
  int *rdone;
};

enum : uint8_t { FLAG_SNH = 1, FLAG_PLASTIC = 2, FLAG_TEAR = 4, FLAG_DAMP = 8 };

// The work item (tile or task) of workgroup b in a grid of n.  Workgroups
// are dealt to the GPU's 8 XCDs round-robin (workgroup b to XCD b % 8), each
// XCD with its own L2; giving XCD x one contiguous run of items keeps what
// neighbouring items share (nodes of adjacent tiles, a front's vector and
// panels across its tasks) in one L2.  A bijection of [0, n).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
  constexpr uint32_t XCDS = 8;
  const uint32_t q = n / XCDS, r = n % XCDS, x = b % XCDS, i = b / XCDS;
  return x * q + min(x, r) + i;
}

// iterate updates of the plastic Newton, on the free nodes of active bodies
enum NewtonVec : int {
  NV_BACKTRACK = 0,  // dx *= 1/2, a -= dx (halve the last increment)
  NV_SAVE = 1,       // lowest = a
  NV_RESTORE = 2,    // a = lowest
  NV_RESET = 3,      // a = carried, x0 = a, x_eval = x_w (acceleration 0 total)
  NV_DX_FROM_A = 4,  // dx = a (the first solve's increment is the whole total)
};

// Adjoint of one converged substep (differentiable rollout): adjoint state,
// accumulated gradients and the recorded substep (x_n, v_n, a_n, a').
struct AdjDev {
  double *xb, *vb, *ab;       // 3N adjoint of x, v, a
  double *w;                  // N, mass-term weight (inertia a' + am a_n - g) . lambda
  double *kl;                 // 3N, K(x_w) lambda (all rows)
  double *lam_bar, *mu_bar;   // T
  double *xslot;              // 12 T, per-corner rest-position gradient
  double *X_bar;              // 3N
  const double *xn, *vn, *an, *anext;
};

template <class Real>
struct Launch {
  static void predict(const Dev<Real> &d, hipStream_t s);
  static void predictor_inverts(const Dev<Real> &d, hipStream_t s);
  static void evaluate(const Dev<Real> &d, hipStream_t s);
  static void assemble(const Dev<Real> &d, bool diag, bool energy, hipStream_t s,
                       int viscous = VISCOUS_NONE);
  static void balance(const Dev<Real> &d, bool diag, bool taken, hipStream_t s);
  static void pcg_init(const Dev<Real> &d, bool with_x0, bool factor, hipStream_t s);
  static void pcg_rz(const Dev<Real> &d, bool start, int parity, hipStream_t s);
  static void pcg_begin(const Dev<Real> &d, hipStream_t s);
  static void op_element(const Dev<Real> &d, const Real *dir, int parity, hipStream_t s);
  static void op_gather(const Dev<Real> &d, const Real *dir, Real *out, int parity, hipStream_t s);
  static void pcg_update(const Dev<Real> &d, int parity, bool factor, hipStream_t s);
  static void pcg_direction(const Dev<Real> &d, int parity, hipStream_t s);
  static void apply_increment(const Dev<Real> &d, bool from_x0, hipStream_t s);
  // FP64 iterative refinement: rres = right - A u with the operator in FP64
  // arithmetic on the FP64 rest data, u the FP64 unknown of the last solve
  // (a, or inc, the FP64 sum of the increments a correction solve added to
  // a), rrel = |rres| / |right| per body, rdone = rrel <= tolerance (and
  // rres zeroed for those bodies, so a correction solve leaves them alone)
  static void refine_residual(const Dev<Real> &d, const double *inc, double tolerance, hipStream_t s);
  // d.rbase = dx (add false) or d.rbase += dx, in FP64
  static void track_increment(const Dev<Real> &d, bool add, hipStream_t s);
  static void finish(const Dev<Real> &d, hipStream_t s);
  static void commit(const Dev<Real> &d, hipStream_t s);
  static void tear(const Dev<Real> &d, hipStream_t s);
  static void kinetic(const Dev<Real> &d, hipStream_t s);
  // kernel-level entry points
  static void displacement(const Dev<Real> &d, const double *xin, hipStream_t s);
  static void apply_plain(const Dev<Real> &d, const Real *dir, Real *out, double inertia,
                          double reach, hipStream_t s);
  static void op_element_plain(const Dev<Real> &d, const Real *dir, hipStream_t s);
  // one body's operator on body-local FP64 device vectors, its 4 T_b corner
  // slots in `slots` (3 T_b * 4 values) instead of the engine's arrays
  static void apply_body(const Dev<Real> &d, uint32_t t0, uint32_t t1, uint32_t n0, uint32_t n1,
                         uint32_t k0, const double *din, double *yout, double inertia, double reach,
                         Real *slots, hipStream_t s);
  // the same with the initial-stress term of corotated tets at positions xin
  // (after an assembly at xin, whose corner forces the slots hold)
  static void apply_body_stressed(const Dev<Real> &d, uint32_t t0, uint32_t t1, uint32_t n0, uint32_t n1,
                                  uint32_t k0, const double *xin, const double *din, double *yout,
                                  double inertia, double reach, Real *slots, hipStream_t s);
  // the same with the damping term: inertia M d + reach K d + damp C d, C
  // the damping operator of the fits of the last assembly
  static void apply_body_damped(const Dev<Real> &d, uint32_t t0, uint32_t t1, uint32_t n0, uint32_t n1,
                                uint32_t k0, const double *din, double *yout, double inertia, double reach,
                                double damp, Real *slots, hipStream_t s);
  static void gather_forces(const Dev<Real> &d, Real *out, hipStream_t s);
  // convergence of the corrections: per node tile max |dx| and max |a|
  static void increment_norms(const Dev<Real> &d, double *out, hipStream_t s);
  // plastic Newton: per node tile |right|^2, max node |dx|, max node
  // |x_eval| (d.nnorm); and the iterate updates of NewtonVec
  static void newton_norms(const Dev<Real> &d, hipStream_t s);
  // bodyRotation (factor.cc:615-646): the positions of the refreshed bodies
  // become their factor's reference; the frame of every active body is the
  // polar factor of its mass-weighted correlation of positions now and at
  // the factorization; F^T r for the bodies still iterating
  static void frame_reference(const Dev<Real> &d, const int *refresh, hipStream_t s);
  static void frames(const Dev<Real> &d, hipStream_t s);
  static void frame_in(const Dev<Real> &d, int parity, hipStream_t s);
  static void newton_vec(const Dev<Real> &d, int mode, hipStream_t s);
  // adjoint of a converged substep (neo-Hookean Hessian)
  static void hvp(const Dev<Real> &d, const Real *dir, int parity, hipStream_t s);
  static void adj_prepare(const Dev<Real> &d, const AdjDev &a, hipStream_t s);
  static void adj_update(const Dev<Real> &d, const AdjDev &a, hipStream_t s);
  static void adj_params(const Dev<Real> &d, const AdjDev &a, hipStream_t s);
};

}  // namespace tmi
