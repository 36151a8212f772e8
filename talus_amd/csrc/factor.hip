// factor.hip -- gfx950 kernels of the supernodal block Cholesky
// preconditioner: FP64 assembly of the explicit system, extend-add, blocked
// front factorization (POTRF / TRSM / SYRK on 64x64 tiles), inversion of
// the diagonal blocks, and the level-by-level triangular solve.
//
// Fronts are dense column-major m x m FP64 matrices (ld = m) in one pool;
// only the lower triangle is meaningful.  After factorization the first k
// columns hold [L11; L21] and the trailing (m-k)^2 block the update matrix
// passed to the parent.  M11 = L11^-1 is stored per front (k x k, ld = k).
#include <cfloat>

#include "factor.h"
#include "kernels.h"
#include "tmath.h"

namespace tmi {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;  // threads per workgroup
static_assert(GCHUNK == NT, "one gather thread per front row of a chunk");
constexpr int LDP = FNB + 1;

__device__ __forceinline__ uint32_t lo32(uint64_t t) { return uint32_t(t & 0xffffffffu); }
__device__ __forceinline__ uint32_t mid16(uint64_t t) { return uint32_t((t >> 32) & 0xffffu); }
__device__ __forceinline__ uint32_t hi16(uint64_t t) { return uint32_t(t >> 48); }

// ------------------------------------------------------------- assembly

// One lower node block of A = inertia M + h^2 sum_e R K_e R^T per thread,
// summed over its tets in ascending order (factor.cc:32-88 precise blocks);
// neo-Hookean tets contribute the operator's blocks (shear part rotated,
// volumetric part along the volume frame W).
template <class Real>
__global__ __launch_bounds__(NT) void k_assemble_pairs(
    double *pool, const uint64_t *pair_target, const uint32_t *pair_ld, const uint32_t *pair_node,
    const uint32_t *pair_start, const uint32_t *contrib, uint32_t pairs, const double *dminv,
    const double *vol, const double *lam, const double *mu, const Real *R, const Real *W,
    const double *plastic, const uint8_t *flags, uint32_t T, const double *mass, const double *reach_of,
    const double *inertia_of, const uint32_t *pair_body, const int *refresh, FactorDamping<Real> damping) {
  const uint32_t q = blockIdx.x * NT + threadIdx.x;
  if (q >= pairs || !refresh[pair_body[q]]) return;
  double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double Bd[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // sum tau C_ab of the damped tets
  const double damp = damping.tau ? damping.body_damp[pair_body[q]] : 0.0;
  for (uint32_t c = pair_start[q]; c < pair_start[q + 1]; ++c) {
    const uint32_t code = contrib[c];
    const uint32_t e = code >> 4, a = (code >> 2) & 3u, b = code & 3u;
    double dm[9], g[12], Rm[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      dm[k] = dminv[size_t(k) * T + e];
      Rm[k] = double(R[size_t(k) * T + e]);
    }
    shape_gradients(dm, g);
    const bool snh = (flags[e] & FLAG_SNH) != 0;
    // the shear part of a plastic neo-Hookean tet is taken along its plastic
    // reference's gradients Fp^-T g = (I + ep)^T g, as in the operator
    double gs[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) gs[i] = g[i];
    if (snh && (flags[e] & FLAG_PLASTIC)) {
      double ep[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) ep[k] = plastic[size_t(k) * T + e];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 3; ++r)
          gs[3 * i + r] += ep[r] * g[3 * i] + ep[3 + r] * g[3 * i + 1] + ep[6 + r] * g[3 * i + 2];
    }
    const double *ga = &gs[3 * a], *gb = &gs[3 * b];
    const double V = vol[e], l = snh ? 0.0 : lam[e], m = mu[e];
    const double gg = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
    double K[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int s = 0; s < 3; ++s)
        K[r * 3 + s] = V * (l * ga[r] * gb[s] + m * gb[r] * ga[s] + (r == s ? m * gg : 0.0));
    double RK[9], RKR[9];
    mul3(Rm, K, RK);
    mul3_nt(RK, Rm, RKR);
#pragma unroll
    for (int k = 0; k < 9; ++k) B[k] += RKR[k];
    if (snh) {
      // volumetric part along the volume frame: V lambda (W g_a)(W g_b)^T
      double Wm[9], wa[3], wb[3];
#pragma unroll
      for (int k = 0; k < 9; ++k) Wm[k] = double(W[size_t(k) * T + e]);
      mv3(Wm, &g[3 * a], wa);
      mv3(Wm, &g[3 * b], wb);
      const double vl = V * lam[e];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) B[r * 3 + s] += vl * wa[r] * wb[s];
    }
    if (damp != 0.0 && (flags[e] & FLAG_DAMP)) {
      // block (a, b) of tau C: column s is corner a of C e_(b, s)
      // (dampingBlocks, factor.cc:273-297), with the rest gradients
      double xc[12], fit[FIT_PLANES];
      const uint4 c4 = damping.conn[e];
      const uint32_t nn[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 3; ++r) xc[3 * i + r] = damping.x[3 * size_t(nn[i]) + r];
      rigid_fit(xc, fit);
      const double tau = double(damping.tau[e]);
      for (int s = 0; s < 3; ++s) {
        double unit[12], col[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) unit[k] = k == int(3 * b) + s ? 1.0 : 0.0;
        damp_tet(fit, Rm, g, V, lam[e], m, unit, col);
#pragma unroll
        for (int r = 0; r < 3; ++r) Bd[r * 3 + s] += tau * col[3 * a + r];
      }
    }
  }
  const uint32_t body = pair_body[q];
  const double reach = reach_of[body];
#pragma unroll
  for (int k = 0; k < 9; ++k) B[k] = reach * B[k] + damp * Bd[k];
  if (pair_node[q] != ~0u) {
    const double im = inertia_of[body] * mass[pair_node[q]];
    B[0] += im;
    B[4] += im;
    B[8] += im;
  }
  double *o = pool + pair_target[q];
  const uint64_t ld = pair_ld[q];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int s = 0; s < 3; ++s) o[uint64_t(s) * ld + r] = B[r * 3 + s];
}

// true when this factorization rebuilds the front's body
__device__ __forceinline__ bool stale(const FactorDev &f, uint32_t fr) {
  return f.refresh[f.body[fr]] != 0;
}

__global__ __launch_bounds__(NT) void k_zero(FactorDev f, const uint64_t *tasks) {
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), chunk = mid16(t);
  if (!stale(f, fr)) return;
  const bool m11 = hi16(t) != 0;
  const uint64_t n = m11 ? uint64_t(f.k[fr]) * f.k[fr] : uint64_t(f.m[fr]) * f.m[fr];
  double *base = m11 ? f.minv + f.minv_off[fr] : f.pool + f.off[fr];
  const uint64_t b0 = uint64_t(chunk) * ZCHUNK, b1 = min(n, b0 + ZCHUNK);
  for (uint64_t i = b0 + threadIdx.x; i < b1; i += NT) base[i] = 0.0;
}

// --------------------------------------------------------------- extend-add

__global__ __launch_bounds__(NT) void k_extend_add(FactorDev f, const uint32_t *parent,
                                                   const uint64_t *tasks) {
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t c = lo32(t), ti = mid16(t), tj = hi16(t);
  if (!stale(f, c)) return;
  const uint32_t p = parent[c];
  const uint32_t kc = f.k[c], mc = f.m[c], mp = f.m[p];
  const uint32_t u = mc - kc;
  const double *src = f.pool + f.off[c];
  double *dst = f.pool + f.off[p];
  const uint32_t *map = f.map + f.map_off[c];
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const uint32_t r = ti * FNB + (idx % FNB), q = tj * FNB + (idx / FNB);
    if (r >= u || q >= u || r < q) continue;
    const uint32_t pr = 3 * map[r / 3] + r % 3, pc = 3 * map[q / 3] + q % 3;
    dst[uint64_t(pc) * mp + pr] += src[uint64_t(kc + q) * mc + kc + r];
  }
}

// -------------------------------------------------------------------- POTRF

// Cholesky of the 64x64 diagonal block of panel p by one wavefront; writes
// L_pp to the front and L_pp^-1 to the (p,p) block of M11.  Lane i keeps row
// i of the block in registers (r[l] = A[i][l], l <= i) and every loop is
// unrolled so the registers are indexed at compile time:
//   factorization (right-looking): for each column j, the pivot comes from
//   lane j by a shuffle, lanes below scale their r[j], and lane i updates
//   r[l] -= r[j] L[l][j] for l > j with L[l][j] shuffled from lane l (lanes
//   above the diagonal compute values they never use);
//   inverse: L is stored to LDS, then lane c forms column c of L^-1 by
//   forward substitution, x[q] in registers and L[i][q] read from LDS (the
//   same address for every lane: a broadcast).
__global__ __launch_bounds__(64) void k_potrf(FactorDev f, const uint32_t *tasks, uint32_t panel) {
  __shared__ double Ls[FNB][LDP];
  const uint32_t fr = tasks[blockIdx.x];
  if (!stale(f, fr)) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const uint32_t P0 = panel * FNB;
  const int nb = int(min(uint32_t(FNB), k - P0));
  double *A = f.pool + f.off[fr];
  const int i = threadIdx.x;
  double r[FNB];
#pragma unroll
  for (int l = 0; l < FNB; ++l)
    r[l] = (i < nb && l < nb && l <= i) ? A[uint64_t(P0 + l) * m + P0 + i] : (l == i ? 1.0 : 0.0);
  bool bad = false;
#pragma unroll
  for (int j = 0; j < FNB; ++j) {
    if (j < nb) {
      double d = __shfl(r[j], j, 64);
      if (!(d > 0.0)) {
        bad = true;
        d = 1.0;
      }
      const double piv = sqrt(d);
      if (i == j) r[j] = piv;
      if (i > j) r[j] /= piv;
#pragma unroll
      for (int l = j + 1; l < FNB; ++l) {
        const double lj = __shfl(r[j], l, 64);  // L[l][j]
        if (l <= i) r[l] -= r[j] * lj;
      }
    }
  }
  if (bad && i == 0) f.failed[f.body[fr]] = 1;
#pragma unroll
  for (int l = 0; l < FNB; ++l) Ls[i][l] = l <= i ? r[l] : 0.0;
  if (i < nb) {
#pragma unroll
    for (int l = 0; l < FNB; ++l)
      if (l <= i && l < nb) A[uint64_t(P0 + l) * m + P0 + i] = r[l];
  }
  __syncthreads();
  // column c = i of L^-1: x[q] = 0 above the diagonal, 1 / L[c][c] on it,
  // -(sum_{c<=q<t} L[t][q] x[q]) / L[t][t] below
  double x[FNB];
#pragma unroll
  for (int t = 0; t < FNB; ++t) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < t; ++q)
      if (q >= i) s += Ls[t][q] * x[q];
    x[t] = t < i ? 0.0 : (t == i ? 1.0 : -s) / Ls[t][t];
  }
  double *M = f.minv + f.minv_off[fr];
  // M11 column i holds x; lanes write row t of the block together
  __syncthreads();
#pragma unroll
  for (int t = 0; t < FNB; ++t) Ls[t][i] = x[t];
  __syncthreads();
  for (int c = 0; c < nb; ++c)
    if (i >= c && i < nb) M[uint64_t(P0 + c) * k + P0 + i] = Ls[i][c];
}

// --------------------------------------------------------- 64x64 tile GEMM

// FP64 MFMA tile: C[r][c] += sum_{kk<64} As[kk][r] Bs[kk][c] for the 64x64
// tile, LDS tiles zero-padded to 64.  Four waves, each a 32x32 quadrant of
// 2x2 v_mfma_f64_16x16x4f64 blocks.  Operand layout of that instruction:
// lane l holds A[l%16][l/16] and B[l/16][l%16]; result item v of lane l is
// D[4v + l/16][l%16].
struct MfmaAcc {
  d4 c[2][2];
};

__device__ __forceinline__ void mfma_zero(MfmaAcc &acc) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc.c[a][b] = d4{0.0, 0.0, 0.0, 0.0};
}

__device__ __forceinline__ void mfma_tile(const double (*As)[LDP], const double (*Bs)[LDP],
                                          MfmaAcc &acc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int qr = 32 * (wave >> 1), qc = 32 * (wave & 1);
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int k0 = 0; k0 < FNB; k0 += 4) {
    const double a0 = As[k0 + lk][qr + li], a1 = As[k0 + lk][qr + 16 + li];
    const double b0 = Bs[k0 + lk][qc + li], b1 = Bs[k0 + lk][qc + 16 + li];
    acc.c[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc.c[0][0], 0, 0, 0);
    acc.c[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc.c[0][1], 0, 0, 0);
    acc.c[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc.c[1][0], 0, 0, 0);
    acc.c[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc.c[1][1], 0, 0, 0);
  }
}

// writes the accumulated tile to Cs[r][c] (LDS); caller synchronizes
__device__ __forceinline__ void mfma_store(const MfmaAcc &acc, double (*Cs)[LDP]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int qr = 32 * (wave >> 1), qc = 32 * (wave & 1);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        Cs[qr + 16 * a + 4 * v + (lane >> 4)][qc + 16 * b + (lane & 15)] = acc.c[a][b][v];
}

// TRSM: L_{rt,p} = A_{rt,p} L_pp^-T for the rows below the panel (all m
// rows; tile rt starts 64 rt rows after the panel's last row).
__global__ __launch_bounds__(NT) void k_trsm(FactorDev f, const uint64_t *tasks, uint32_t panel) {
  __shared__ double As[FNB][LDP];  // As[q][r] = A[R0+r][P0+q]
  __shared__ double Xs[FNB][LDP];  // Xs[q][c] = X[c][q] = (L_pp^-1)[c][q]
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), rt = mid16(t);
  if (!stale(f, fr)) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const uint32_t P0 = panel * FNB;
  const int nb = int(min(uint32_t(FNB), k - P0));
  const uint32_t R0 = P0 + nb + rt * FNB;  // trailing rows start after the panel
  const int nr = int(min(uint32_t(FNB), m - R0));
  double *A = f.pool + f.off[fr];
  const double *X = f.minv + f.minv_off[fr];
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int r = idx % FNB, q = idx / FNB;
    As[q][r] = (r < nr && q < nb) ? A[uint64_t(P0 + q) * m + R0 + r] : 0.0;
    // X column q rows r: X[r][q] lower (r >= q)
    Xs[q][r] = (r < nb && q < nb && r >= q) ? X[uint64_t(P0 + q) * k + P0 + r] : 0.0;
  }
  __syncthreads();
  MfmaAcc acc;
  mfma_zero(acc);
  mfma_tile(As, Xs, acc);
  __syncthreads();
  mfma_store(acc, As);  // As reused as C[r][c]
  __syncthreads();
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int r = idx % FNB, c = idx / FNB;
    if (r < nr && c < nb) A[uint64_t(P0 + c) * m + R0 + r] = As[r][c];
  }
}

// SYRK/GEMM trailing update A_IJ -= L_Ip L_Jp^T over the rows/columns after
// the panel (tiles counted from the panel's last row).
__global__ __launch_bounds__(NT) void k_syrk(FactorDev f, const uint64_t *tasks, uint32_t panel) {
  __shared__ double Is[FNB][LDP];
  __shared__ double Js[FNB][LDP];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), I = mid16(t), J = hi16(t);
  if (!stale(f, fr)) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const uint32_t P0 = panel * FNB;
  const int nb = int(min(uint32_t(FNB), k - P0));
  const uint32_t RI = P0 + nb + I * FNB, RJ = P0 + nb + J * FNB;
  const int ni = int(min(uint32_t(FNB), m - RI)), nj = int(min(uint32_t(FNB), m - RJ));
  double *A = f.pool + f.off[fr];
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int r = idx % FNB, q = idx / FNB;
    const bool inq = q < nb;
    Is[q][r] = (inq && r < ni) ? A[uint64_t(P0 + q) * m + RI + r] : 0.0;
    Js[q][r] = (inq && r < nj) ? A[uint64_t(P0 + q) * m + RJ + r] : 0.0;
  }
  __syncthreads();
  MfmaAcc acc;
  mfma_zero(acc);
  mfma_tile(Is, Js, acc);
  __syncthreads();
  mfma_store(acc, Is);  // Is reused as C[r][c]
  __syncthreads();
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int r = idx % FNB, c = idx / FNB;
    if (r < ni && c < nj) A[uint64_t(RJ + c) * m + RI + r] -= Is[r][c];
  }
}

// M11 = L11^-1 by forward substitution on column blocks: for row block i
// and column block j < i, X_ij = -L_ii^-1 sum_{q=j}^{i-1} L_iq X_qj.
__global__ __launch_bounds__(NT) void k_trtri(FactorDev f, const uint64_t *tasks) {
  __shared__ double As[FNB][LDP];  // As[q][r] = L[i-block r][q-block col]
  __shared__ double Bs[FNB][LDP];  // Bs[q][c] = X[q-block row][j-block c]
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), i = mid16(t), j = hi16(t);
  if (!stale(f, fr)) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const double *L = f.pool + f.off[fr];
  double *X = f.minv + f.minv_off[fr];
  const int ni = int(min(uint32_t(FNB), k - i * FNB)), nj = int(min(uint32_t(FNB), k - j * FNB));
  MfmaAcc acc;
  mfma_zero(acc);
  for (uint32_t q = j; q < i; ++q) {
    const int nq = int(min(uint32_t(FNB), k - q * FNB));
    __syncthreads();
    for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
      const int r = idx % FNB, c = idx / FNB;
      As[c][r] = (r < ni && c < nq) ? L[uint64_t(q * FNB + c) * m + i * FNB + r] : 0.0;
      // X[q*FNB + r][j*FNB + c]: lower part of M11 (row >= col)
      const uint32_t xr = q * FNB + r, xc = j * FNB + c;
      Bs[r][c] = (r < nq && c < nj && xr >= xc) ? X[uint64_t(xc) * k + xr] : 0.0;
    }
    __syncthreads();
    mfma_tile(As, Bs, acc);
  }
  // X_ij = -Linv_ii * S with S the accumulated sum
  __syncthreads();
  mfma_store(acc, Bs);  // Bs[r][c] = S[r][c]
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int r = idx % FNB, c = idx / FNB;
    // As[c][r] = Linv_ii[r][c]
    const uint32_t xr = i * FNB + r, xc = i * FNB + c;
    As[c][r] = (r < ni && c < ni && r >= c) ? X[uint64_t(xc) * k + xr] : 0.0;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int r = idx % FNB, c = idx / FNB;
    if (r >= ni || c >= nj) continue;
    double s = 0.0;
    for (int q = 0; q <= r; ++q) s += As[q][r] * Bs[q][c];
    X[uint64_t(j * FNB + c) * k + i * FNB + r] = -s;
  }
}

// ----------------------------------------------------------------- solve

__device__ __forceinline__ bool skip(const int *done, int parity, uint32_t B, uint32_t body) {
  return parity >= 0 && done[parity * B + body];
}

// A front's sizes and offsets, loaded together before the convergence test
// (loads after the test's branch would wait for it)
struct FrontMeta {
  uint32_t k, m, body, rows_off;
  uint64_t g_off, yoff, vec_off;
  __device__ __forceinline__ FrontMeta(const FactorDev &f, uint32_t fr)
      : k(f.k[fr]), m(f.m[fr]), body(f.body[fr]), rows_off(f.rows_off[fr]), g_off(f.g_off[fr]),
        yoff(f.yoff[fr]), vec_off(f.vec_off[fr]) {}
};

// forward gather of row i of a front: r at the front's columns (i < k) plus
// the children's update entries that map to row i, pulled in child order
template <class Real>
__device__ __forceinline__ double gather_row(const FactorDev &f, uint32_t rows_off, uint32_t i, uint32_t k,
                                             const Real *rin) {
  const uint32_t nrow = rows_off + i / 3, d = i % 3;
  double v = i < k ? double(rin[3 * size_t(f.rows[nrow]) + d]) : 0.0;
  for (uint32_t q = f.src_start[nrow]; q < f.src_start[nrow + 1]; ++q) v += f.vec[f.src[q] + d];
  return v;
}

// the rows shape's front vectors, gathered once per front: v[i] for the
// chunk's rows (forward), v[i] = x2 at the update rows (backward)
template <class Real>
__global__ __launch_bounds__(NT) void k_solve_gather(FactorDev f, const uint64_t *tasks, const Real *rin,
                                                     const int *done, int parity, uint32_t B) {
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), chunk = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const uint32_t i = chunk * GCHUNK + threadIdx.x;
  if (i >= m) return;
  f.vec[f.vec_off[fr] + i] = gather_row(f, f.rows_off[fr], i, k, rin);
}

__global__ __launch_bounds__(NT) void k_solve_xgather(FactorDev f, const uint64_t *tasks, const int *done,
                                                      int parity, uint32_t B) {
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), chunk = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const uint32_t i = k + chunk * GCHUNK + threadIdx.x;
  if (i >= m) return;
  const uint32_t *rows = f.rows + f.rows_off[fr];
  f.vec[f.vec_off[fr] + i] = f.xsol[3 * size_t(rows[i / 3]) + i % 3];
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// stored entries of G and their FP64 value
__device__ __forceinline__ double gval(double x) { return x; }
__device__ __forceinline__ double gval(float x) { return double(x); }
__device__ __forceinline__ double gval(__half x) { return double(__half2float(x)); }
template <class G>
__device__ __forceinline__ G gput(double x);
template <>
__device__ __forceinline__ double gput<double>(double x) { return x; }
template <>
__device__ __forceinline__ float gput<float>(double x) { return float(x); }
template <>
__device__ __forceinline__ __half gput<__half>(double x) {
  // saturate instead of overflowing to inf (FP16 max 65504)
  return __float2half(fminf(fmaxf(float(x), -65504.0f), 65504.0f));
}
// s_j of column j: the power of two nearest the pivot inverse M11[j][j] > 0,
// so G[i][j] / s_j is exact and the stored column is O(1)
__device__ __forceinline__ double col_scale(double diag) {
  return diag > 0.0 && diag < DBL_MAX ? ldexp(1.0, int(rint(log2(diag)))) : 1.0;
}

// G's two copies are tiled by W = 16 / sizeof(G) entries so that a lane
// loads 16 bytes at a time and a wavefront's 64 lanes read one contiguous
// kilobyte:
//   column-tiled (gc): entry (i, j) at (j / W * m + i) * W + j % W -- the W
//     consecutive columns of a row are contiguous, consecutive rows adjacent;
//   row-tiled (gr): entry (i, j) at (i / W * k + j) * W + i % W -- the W
//     consecutive rows of a column are contiguous, consecutive columns
//     adjacent.
// Column tiles past k and row tiles past m are padded with zeros (the pool
// is cleared when allocated and the padding is never written).
template <class G>
struct Tile {
  static constexpr uint32_t W = 16u / uint32_t(sizeof(G));
};
template <class G>
struct alignas(16) Chunk {
  G v[16 / sizeof(G)];
};
template <class G>
__device__ __forceinline__ Chunk<G> load_chunk(const G *p) {
  return *reinterpret_cast<const Chunk<G> *>(p);
}
__device__ __forceinline__ uint64_t gc_at(uint32_t i, uint32_t j, uint32_t m, uint32_t W) {
  return (uint64_t(j / W) * m + i) * W + j % W;
}
__device__ __forceinline__ uint64_t gr_at(uint32_t i, uint32_t j, uint32_t k, uint32_t W) {
  return (uint64_t(i / W) * k + j) * W + i % W;
}

// acc[q % 4] += g[b * ts + q] * w[b * W + q] over tiles b in [b0, b1) and
// q < W: one lane's dot product along its row (gc) or column (gr), four
// 16-byte tiles in flight
template <class G>
__device__ __forceinline__ void tile_dot(const G *g, uint64_t ts, const double *w, uint32_t b0,
                                         uint32_t b1, double *acc) {
  constexpr uint32_t W = Tile<G>::W, UB = 4;
  uint32_t b = b0;
  for (; b + UB <= b1; b += UB) {
    Chunk<G> c[UB];
#pragma unroll
    for (uint32_t u = 0; u < UB; ++u) c[u] = load_chunk(g + uint64_t(b + u) * ts);
#pragma unroll
    for (uint32_t u = 0; u < UB; ++u)
#pragma unroll
      for (uint32_t q = 0; q < W; ++q) acc[q % 4] += gval(c[u].v[q]) * w[(b + u) * W + q];
  }
  for (; b < b1; ++b) {
    const Chunk<G> c = load_chunk(g + uint64_t(b) * ts);
#pragma unroll
    for (uint32_t q = 0; q < W; ++q) acc[q % 4] += gval(c.v[q]) * w[b * W + q];
  }
}

// The first PF tiles of a lane's run [b0, b1), loaded before the front
// vector staged in LDS is needed: G does not depend on the vector, so these
// loads stay in flight across the barrier (which waits for LDS, not for
// outstanding global loads).  The loads are unconditional, clamped to
// `last`, a tile of the lane's row or column that exists (a conditional load
// would make the compiler wait for it where the branches join); dot() then
// adds exactly what tile_dot(b0, b1) adds, in the same order.
#ifndef TMI_PF
#define TMI_PF 8
#endif
constexpr uint32_t PF = TMI_PF;
template <class G>
struct TilePrefetch {
  Chunk<G> c[PF];
  __device__ __forceinline__ void issue(const G *g, uint64_t ts, uint32_t b0, uint32_t last) {
#pragma unroll
    for (uint32_t u = 0; u < PF; ++u) c[u] = load_chunk(g + uint64_t(min(b0 + u, last)) * ts);
  }
  // the same for a lane whose tiles outside [lo, hi] lie above the diagonal
  // of M11 (row-tiled copy, lane = column j: b < lo = j / W; column-tiled
  // copy, lane = row i < k: b > hi = i / W).  Those hold zeros and are read
  // at tile zt instead, which the caller picks among the lane's zero tiles
  // whenever there are any, so the lane sums the same values while the
  // repeats of one tile hit in cache instead of fetching zeros from HBM
  __device__ __forceinline__ void issue(const G *g, uint64_t ts, uint32_t b0, uint32_t last, uint32_t lo,
                                        uint32_t hi, uint32_t zt) {
#pragma unroll
    for (uint32_t u = 0; u < PF; ++u) {
      const uint32_t b = b0 + u;
      c[u] = load_chunk(g + uint64_t(b < lo || b > hi ? zt : min(b, last)) * ts);
    }
  }
  __device__ __forceinline__ void dot(const G *g, uint64_t ts, const double *w, uint32_t b0, uint32_t b1,
                                      double *acc) const {
    constexpr uint32_t W = Tile<G>::W;
#pragma unroll
    for (uint32_t u = 0; u < PF; ++u)
      if (b0 + u < b1)
#pragma unroll
        for (uint32_t q = 0; q < W; ++q) acc[q % 4] += gval(c[u].v[q]) * w[(b0 + u) * W + q];
    if (b0 + PF < b1) tile_dot(g, ts, w, b0 + PF, b1, acc);
  }
};

// G = [M11; W21]: W21 tile (I, J) = sum_{q >= J} L21[I][q] M11[q][J]
template <class Real>
__global__ __launch_bounds__(NT) void k_pack_w21(FactorDev f, const uint64_t *tasks,
                                                 typename GStore<Real>::type *gr,
                                                 typename GStore<Real>::type *gc) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W;
  __shared__ double As[FNB][LDP];  // As[kk][r] = L21[I*64 + r][q*64 + kk]
  __shared__ double Bs[FNB][LDP];  // Bs[kk][c] = M11[q*64 + kk][J*64 + c]
  __shared__ double inv_s[FNB];    // 1 / s_j of the tile's columns
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), I = mid16(t), J = hi16(t);
  if (!stale(f, fr)) return;
  const uint32_t k = f.k[fr], m = f.m[fr], u = m - k;
  const double *L = f.pool + f.off[fr];
  const double *M = f.minv + f.minv_off[fr];
  const int ni = int(min(uint32_t(FNB), u - I * FNB)), nj = int(min(uint32_t(FNB), k - J * FNB));
  MfmaAcc acc;
  mfma_zero(acc);
  const uint32_t kt = (k + FNB - 1) / FNB;
  for (uint32_t q = J; q < kt; ++q) {
    const int nq = int(min(uint32_t(FNB), k - q * FNB));
    __syncthreads();
    for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
      const int r = idx % FNB, c = idx / FNB;
      As[c][r] = (r < ni && c < nq) ? L[uint64_t(q * FNB + c) * m + k + I * FNB + r] : 0.0;
      const uint32_t mr = q * FNB + r, mc = J * FNB + c;
      Bs[r][c] = (r < nq && c < nj && mr >= mc) ? M[uint64_t(mc) * k + mr] : 0.0;
    }
    __syncthreads();
    mfma_tile(As, Bs, acc);
  }
  __syncthreads();
  mfma_store(acc, As);  // As[r][c] = W21 tile
  if (threadIdx.x < FNB && int(threadIdx.x) < nj) {
    const uint32_t gj = J * FNB + threadIdx.x;
    inv_s[threadIdx.x] = 1.0 / col_scale(M[uint64_t(gj) * k + gj]);
  }
  __syncthreads();
  G *R = gr + f.g_off[fr];
  G *C = gc + f.g_off[fr];
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int r = idx % FNB, c = idx / FNB;
    if (r >= ni || c >= nj) continue;
    const uint32_t gi = k + I * FNB + r, gj = J * FNB + c;
    const G v = gput<G>(As[r][c] * inv_s[c]);
    C[gc_at(gi, gj, m, W)] = v;
    R[gr_at(gi, gj, k, W)] = v;
  }
}

// G rows 0..k-1 = M11 (lower tiles; the upper part is stored as zeros)
template <class Real>
__global__ __launch_bounds__(NT) void k_pack_m11(FactorDev f, const uint64_t *tasks,
                                                 typename GStore<Real>::type *gr,
                                                 typename GStore<Real>::type *gc) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W;
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), I = mid16(t), J = hi16(t);
  if (!stale(f, fr)) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const double *M = f.minv + f.minv_off[fr];
  G *R = gr + f.g_off[fr];
  G *C = gc + f.g_off[fr];
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const uint32_t i = I * FNB + idx % FNB, j = J * FNB + idx / FNB;
    if (i >= k || j >= k) continue;
    const double sj = col_scale(M[uint64_t(j) * k + j]);
    if (i == j) f.gscale[f.yoff[fr] + j] = sj;
    const G v = i >= j ? gput<G>(M[uint64_t(j) * k + i] / sj) : gput<G>(0.0);
    R[gr_at(i, j, k, W)] = v;
    C[gc_at(i, j, m, W)] = v;
  }
}

// Rows kernels (fronts too large for the lane shapes, and the plan's
// rows-everywhere choice): a workgroup takes RW = 8 rows (forward) or
// columns (backward) of G, i.e. 8 / W tiles, its four wavefronts split the
// other dimension in contiguous quarters, and each lane runs along its
// quarter loading the 16-byte tiles that hold its column's (row's) entries
// for the 8 rows (columns); 8 FP64 wavefront reductions, then the four
// wavefronts' sums added in LDS in wavefront order.  The front vector comes
// from the gather kernels (k_solve_gather / k_solve_xgather), read where it
// lies.
constexpr uint32_t RW = ROWS_GROUP;
static_assert(RW % 8 == 0, "whole tiles of every storage type");

// forward, rows shape: y[i] (i < k) or v[i] - . (i >= k) for the group's
// rows, row-tiled copy
template <class Real>
__global__ __launch_bounds__(NT) void k_solve_forward(FactorDev f, const typename GStore<Real>::type *gr,
                                                      const uint64_t *tasks, const Real *rin,
                                                      const int *done, int parity, uint32_t B) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W, NTL = RW / W;
  (void)rin;  // the front vector was gathered by k_solve_gather
  __shared__ double red[NT / 64][RW];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), grp = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i0 = grp * RW;  // the group's first row (tile aligned)
  const uint32_t ilast = min(i0 + RW, m) - 1;
  const uint32_t n = ilast < k ? ilast + 1 : k;  // columns any row of the group needs
  const uint32_t quarter = (n + 3u) / 4u;
  const uint32_t j0 = min(n, wave * quarter), j1 = min(n, j0 + quarter);
  double *v = f.vec + f.vec_off[fr];
  const double *sc = f.gscale + f.yoff[fr];
  const G *g = gr + f.g_off[fr] + uint64_t(i0 / W) * k * W;
  const uint64_t ts = uint64_t(k) * W;  // between row tiles
  double acc[RW];
#pragma unroll
  for (uint32_t q = 0; q < RW; ++q) acc[q] = 0.0;
  uint32_t j = j0 + lane;
  for (; j + 64 < j1; j += 128) {  // 2 x 8/W tiles per lane in flight
    Chunk<G> c[2][NTL];
    double w[2];
#pragma unroll
    for (uint32_t u = 0; u < 2; ++u) {
#pragma unroll
      for (uint32_t a = 0; a < NTL; ++a) c[u][a] = load_chunk(g + a * ts + uint64_t(j + 64 * u) * W);
      w[u] = v[j + 64 * u] * sc[j + 64 * u];
    }
#pragma unroll
    for (uint32_t a = 0; a < NTL; ++a)
#pragma unroll
      for (uint32_t q = 0; q < W; ++q) acc[a * W + q] += gval(c[0][a].v[q]) * w[0] + gval(c[1][a].v[q]) * w[1];
  }
  for (; j < j1; j += 64) {
    const double w = v[j] * sc[j];
#pragma unroll
    for (uint32_t a = 0; a < NTL; ++a) {
      const Chunk<G> c = load_chunk(g + a * ts + uint64_t(j) * W);
#pragma unroll
      for (uint32_t q = 0; q < W; ++q) acc[a * W + q] += gval(c.v[q]) * w;
    }
  }
#pragma unroll
  for (uint32_t q = 0; q < RW; ++q) {
    const double s = wave_sum64(acc[q]);
    if (lane == 0) red[wave][q] = s;
  }
  __syncthreads();
  if (threadIdx.x >= RW) return;
  const uint32_t i = i0 + threadIdx.x;
  if (i >= m) return;
  const double tot = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
  if (i < k)
    f.y[f.yoff[fr] + i] = tot;
  else
    v[i] -= tot;
}

// forward for fronts with k <= KLANES: one lane per row of G, along the
// row's column tiles of the column-tiled copy; v[0:k] staged in LDS; no
// cross-lane reduction.  A wavefront whose rows all lie below
// j_end = min(k, first row + 64) skips the zero upper triangle of M11
// beyond it.
template <class Real>
__global__ __launch_bounds__(NT) void k_solve_forward_lanes(FactorDev f, const typename GStore<Real>::type *gc,
                                                            const uint64_t *tasks, const Real *rin,
                                                            const int *done, int parity, uint32_t B) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W;
  __shared__ double vs[KLANES];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), chunk = mid16(t);
  const FrontMeta fm(f, fr);
  if (skip(done, parity, B, fm.body)) return;
  const uint32_t k = fm.k, m = fm.m;
  const uint32_t kp = (k + W - 1) / W * W;
  double *v = f.vec + fm.vec_off;
  const double *sc = f.gscale + fm.yoff;
  const uint32_t first = chunk * NT + (threadIdx.x & ~63u);
  const uint32_t i = chunk * NT + threadIdx.x;
  const uint32_t nb = first < m ? (min(k, first + 64u) + W - 1) / W : 0u;  // column tiles the wavefront needs
  const G *g = gc + fm.g_off + uint64_t(min(i, m - 1)) * W;
  // an update row's own right-hand side
  const double own = (i >= k && i < m) ? gather_row(f, fm.rows_off, i, k, rin) : 0.0;
  for (uint32_t q = threadIdx.x; q < kp; q += NT) vs[q] = q < k ? gather_row(f, fm.rows_off, q, k, rin) * sc[q] : 0.0;
  TilePrefetch<G> pre;
  pre.issue(g, uint64_t(m) * W, 0, (k - 1) / W, 0, i < k ? i / W : (k - 1) / W, (k - 1) / W);
  __syncthreads();
  if (first >= m) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  pre.dot(g, uint64_t(m) * W, vs, 0, nb, acc);
  if (i >= m) return;
  const double tot = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  if (i < k)
    f.y[fm.yoff + i] = tot;
  else
    v[i] = own - tot;
}

// The backward kernels' LDS vector w = [y; -x2] from entry `from` to mp.
// Staged in three parts so the tile prefetch overlaps the vector's loads:
// load() reads the workgroup's first two rounds of entries (every entry of
// most fronts) into registers, the caller issues the prefetch, store()
// writes them (waiting for these loads, not for the tiles issued after
// them: global loads complete in order), and rest() stages what is left.
// The empty asm statements keep the compiler from moving loads across.
__device__ __forceinline__ double w_entry(const double *y, const double *xsol, const uint32_t *rows, uint32_t k,
                                          uint32_t m, uint32_t q) {
  return q < k ? y[q] : (q < m ? -xsol[3 * size_t(rows[q / 3]) + q % 3] : 0.0);
}
struct WStage {
  double w0, w1;
  uint32_t q0, q1;
  uint32_t orow;  // node of the lane's output column j
  double oscale;  // s_j
  __device__ __forceinline__ void load(const double *y, const double *xsol, const uint32_t *rows, uint32_t k,
                                       uint32_t m, uint32_t mp, uint32_t from, uint32_t j, const double *gscale) {
    q0 = from + threadIdx.x;
    q1 = q0 + NT;
    const uint32_t jc = min(j, k - 1);
    orow = rows[jc / 3];
    oscale = gscale[jc];
    // unconditional loads at clamped indices, then selects: no branch, so
    // the four value loads are in flight together
    const uint32_t a0 = min(q0, m - 1), a1 = min(q1, m - 1);
    const uint32_t r0 = rows[a0 / 3], r1 = rows[a1 / 3];
    const double y0 = y[min(q0, k - 1)], y1 = y[min(q1, k - 1)];
    const double x0 = xsol[3 * size_t(r0) + a0 % 3], x1 = xsol[3 * size_t(r1) + a1 % 3];
    asm volatile("" ::: "memory");
    w0 = q0 < k ? y0 : (q0 < m ? -x0 : 0.0);
    w1 = q1 < k ? y1 : (q1 < m ? -x1 : 0.0);
    asm volatile("" ::: "memory");
  }
  __device__ __forceinline__ void store(double *ws, uint32_t mp) const {
    asm volatile("" ::: "memory");
    if (q0 < mp) ws[q0] = w0;
    if (q1 < mp) ws[q1] = w1;
  }
  __device__ __forceinline__ void rest(double *ws, const double *y, const double *xsol, const uint32_t *rows,
                                       uint32_t k, uint32_t m, uint32_t mp) const {
    for (uint32_t q = q1 + NT; q < mp; q += NT) ws[q] = w_entry(y, xsol, rows, k, m, q);
  }
};

// backward for fronts with k <= KLANES and m <= MLANES: one lane per column
// of G, along the column's row tiles of the row-tiled copy from the
// wavefront's first column on; w = [y; -x2] staged in LDS, x2 read from the
// solution of the ancestors' levels (xsol at the front's update rows).
template <class Real>
__global__ __launch_bounds__(NT) void k_solve_backward_lanes(FactorDev f, const typename GStore<Real>::type *gr,
                                                             const uint64_t *tasks, const int *done,
                                                             int parity, uint32_t B) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W;
  __shared__ double ws[MLANES];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), chunk = mid16(t);
  const FrontMeta fm(f, fr);
  if (skip(done, parity, B, fm.body)) return;
  const uint32_t k = fm.k, m = fm.m;
  const uint32_t mp = (m + W - 1) / W * W;
  const double *y = f.y + fm.yoff;
  const uint32_t *rows = f.rows + fm.rows_off;
  const uint32_t c0 = chunk * NT;
  const uint32_t first = c0 + (threadIdx.x & ~63u);
  const uint32_t j = c0 + threadIdx.x;
  const G *g = gr + fm.g_off + uint64_t(min(j, k - 1)) * W;
  const uint32_t tb0 = first / W, tb1 = first < k ? mp / W : tb0;
  WStage st;
  st.load(y, f.xsol, rows, k, m, mp, c0, j, f.gscale + fm.yoff);
  TilePrefetch<G> pre;
  pre.issue(g, uint64_t(k) * W, tb0, mp / W - 1, min(j, k - 1) / W, mp / W - 1, tb0);
  st.store(ws, mp);
  st.rest(ws, y, f.xsol, rows, k, m, mp);
  __syncthreads();
  if (first >= k) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  pre.dot(g, uint64_t(k) * W, ws, tb0, tb1, acc);
  if (j >= k) return;
  f.xsol[3 * size_t(st.orow) + j % 3] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) * st.oscale;
}

// forward, split shape (fronts with k <= KLANES): 64 rows of G per
// workgroup, one per lane, the column tiles [0, min(k, first row + 64))
// split in four contiguous quarters over the workgroup's wavefronts
// (column-tiled copy), partial sums added in LDS in wavefront order
template <class Real>
__global__ __launch_bounds__(NT) void k_solve_forward_split(FactorDev f, const typename GStore<Real>::type *gc,
                                                            const uint64_t *tasks, const Real *rin,
                                                            const int *done, int parity, uint32_t B) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W;
  __shared__ double vs[KLANES];
  __shared__ double red[NT / 64][64];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), chunk = mid16(t);
  const FrontMeta fm(f, fr);
  if (skip(done, parity, B, fm.body)) return;
  const uint32_t k = fm.k, m = fm.m;
  double *v = f.vec + fm.vec_off;
  const uint32_t first = chunk * 64u;
  const uint32_t nb = (min(k, first + 64u) + W - 1) / W;
  const double *sc = f.gscale + fm.yoff;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t i = first + lane;
  const uint32_t quarter = (nb + 3u) / 4u;
  const uint32_t b0 = min(nb, wave * quarter), b1 = min(nb, b0 + quarter);
  const G *g = gc + fm.g_off + uint64_t(min(i, m - 1)) * W;
  // an update row's own right-hand side
  const double own = (wave == 0 && i >= k && i < m) ? gather_row(f, fm.rows_off, i, k, rin) : 0.0;
  for (uint32_t q = threadIdx.x; q < nb * W; q += NT) vs[q] = q < k ? gather_row(f, fm.rows_off, q, k, rin) * sc[q] : 0.0;
  TilePrefetch<G> pre;
  pre.issue(g, uint64_t(m) * W, b0, (k - 1) / W, 0, i < k ? i / W : (k - 1) / W, (k - 1) / W);
  __syncthreads();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  pre.dot(g, uint64_t(m) * W, vs, b0, b1, acc);
  red[wave][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (wave != 0 || i >= m) return;
  const double tot = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
  if (i < k)
    f.y[fm.yoff + i] = tot;
  else
    v[i] = own - tot;
}

// backward, split shape (k <= KLANES, m <= MLANES): 64 columns of G per
// workgroup, one per lane, the row tiles [first column, m) split in four
// contiguous quarters over the wavefronts (row-tiled copy), w = [y; -x2]
// staged in LDS, partial sums added in LDS in wavefront order
template <class Real>
__global__ __launch_bounds__(NT) void k_solve_backward_split(FactorDev f, const typename GStore<Real>::type *gr,
                                                             const uint64_t *tasks, const int *done,
                                                             int parity, uint32_t B) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W;
  __shared__ double ws[MLANES];
  __shared__ double red[NT / 64][64];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), chunk = mid16(t);
  const FrontMeta fm(f, fr);
  if (skip(done, parity, B, fm.body)) return;
  const uint32_t k = fm.k, m = fm.m;
  const uint32_t mp = (m + W - 1) / W * W;
  const double *y = f.y + fm.yoff;
  const uint32_t *rowsf = f.rows + fm.rows_off;
  const uint32_t first = chunk * 64u;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t j = first + lane;
  const uint32_t t0 = first / W, tiles = mp / W - t0, quarter = (tiles + 3u) / 4u;
  const uint32_t b0 = t0 + min(tiles, wave * quarter), b1 = t0 + min(tiles, wave * quarter + quarter);
  const G *g = gr + fm.g_off + uint64_t(min(j, k - 1)) * W;
  WStage st;
  st.load(y, f.xsol, rowsf, k, m, mp, first, j, f.gscale + fm.yoff);
  TilePrefetch<G> pre;
  pre.issue(g, uint64_t(k) * W, b0, mp / W - 1, min(j, k - 1) / W, mp / W - 1, t0);
  st.store(ws, mp);
  st.rest(ws, y, f.xsol, rowsf, k, m, mp);
  __syncthreads();
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  pre.dot(g, uint64_t(k) * W, ws, b0, b1, acc);
  red[wave][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (wave != 0 || j >= k) return;
  f.xsol[3 * size_t(st.orow) + j % 3] = (((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]) * st.oscale;
}

// backward, rows shape: x[col j] = sum_{j<=i<k} G[i][j] y[i]
// - sum_{i>=k} G[i][j] x2[i] for the group's columns, column-tiled copy,
// lanes along the rows from the group's first column on (x2 gathered by
// k_solve_xgather)
template <class Real>
__global__ __launch_bounds__(NT) void k_solve_backward(FactorDev f, const typename GStore<Real>::type *gc,
                                                       const uint64_t *tasks, const int *done,
                                                       int parity, uint32_t B) {
  using G = typename GStore<Real>::type;
  constexpr uint32_t W = Tile<G>::W, NTL = RW / W;
  __shared__ double red[NT / 64][RW];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), grp = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t j0 = grp * RW;  // the group's first column (tile aligned)
  const uint32_t rows = m - j0, quarter = (rows + 3u) / 4u;
  const uint32_t r0 = j0 + min(rows, wave * quarter), r1 = j0 + min(rows, wave * quarter + quarter);
  const double *y = f.y + f.yoff[fr];
  const double *v = f.vec + f.vec_off[fr];
  const G *g = gc + f.g_off[fr] + uint64_t(j0 / W) * m * W;
  const uint64_t ts = uint64_t(m) * W;  // between column tiles
  double acc[RW];
#pragma unroll
  for (uint32_t q = 0; q < RW; ++q) acc[q] = 0.0;
  uint32_t i = r0 + lane;
  for (; i + 64 < r1; i += 128) {  // 2 x 8/W tiles per lane in flight
    Chunk<G> c[2][NTL];
    double w[2];
#pragma unroll
    for (uint32_t u = 0; u < 2; ++u) {
      const uint32_t r = i + 64 * u;
#pragma unroll
      for (uint32_t a = 0; a < NTL; ++a) c[u][a] = load_chunk(g + a * ts + uint64_t(r) * W);
      w[u] = r < k ? y[r] : -v[r];
    }
#pragma unroll
    for (uint32_t a = 0; a < NTL; ++a)
#pragma unroll
      for (uint32_t q = 0; q < W; ++q) acc[a * W + q] += gval(c[0][a].v[q]) * w[0] + gval(c[1][a].v[q]) * w[1];
  }
  for (; i < r1; i += 64) {
    const double w = i < k ? y[i] : -v[i];
#pragma unroll
    for (uint32_t a = 0; a < NTL; ++a) {
      const Chunk<G> c = load_chunk(g + a * ts + uint64_t(i) * W);
#pragma unroll
      for (uint32_t q = 0; q < W; ++q) acc[a * W + q] += gval(c.v[q]) * w;
    }
  }
#pragma unroll
  for (uint32_t q = 0; q < RW; ++q) {
    const double s = wave_sum64(acc[q]);
    if (lane == 0) red[wave][q] = s;
  }
  __syncthreads();
  if (threadIdx.x >= RW) return;
  const uint32_t j = j0 + threadIdx.x;
  if (j >= k) return;
  const double tot = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
  const uint32_t *rowsf = f.rows + f.rows_off[fr];
  f.xsol[3 * size_t(rowsf[j / 3]) + j % 3] = tot * f.gscale[f.yoff[fr] + j];
}

template <class Real>
// z = F x for the free nodes (F the body frame of bodyRotation, or the
// identity when frame is null), 0 for anchored ones
__global__ __launch_bounds__(NT) void k_solve_finish(FactorDev f, Real *z, uint32_t N,
                                                     const uint8_t *anchored,
                                                     const uint32_t *node_body, const int *done,
                                                     int parity, uint32_t B, const double *frame) {
  const uint32_t n = blockIdx.x * NT + threadIdx.x;
  if (n >= N || skip(done, parity, B, node_body[n])) return;
  const double *x = f.xsol + 3 * size_t(n);
  if (anchored[n]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) z[3 * size_t(n) + c] = Real(0);
    return;
  }
  if (!frame) {
#pragma unroll
    for (int c = 0; c < 3; ++c) z[3 * size_t(n) + c] = Real(x[c]);
    return;
  }
  const double *F = frame + 9 * size_t(node_body[n]);
#pragma unroll
  for (int c = 0; c < 3; ++c) z[3 * size_t(n) + c] = Real(F[3 * c] * x[0] + F[3 * c + 1] * x[1] + F[3 * c + 2] * x[2]);
}

}  // namespace

// ---------------------------------------------------------------- launchers

template <class Real>
void FactorLaunch<Real>::assemble(const FactorDev &f, const uint64_t *pair_target,
                                  const uint32_t *pair_ld, const uint32_t *pair_node,
                                  const uint32_t *pair_start, const uint32_t *contrib,
                                  uint32_t pairs, const double *dminv, const double *vol,
                                  const double *lam, const double *mu, const Real *R,
                                  const Real *W, const double *plastic, const uint8_t *flags,
                                  uint32_t T, const double *mass, const double *body_reach,
                                  const double *body_inertia, const uint32_t *pair_body,
                                  const FactorDamping<Real> &damping, hipStream_t s) {
  if (!pairs) return;
  hipLaunchKernelGGL(k_assemble_pairs<Real>, dim3((pairs + NT - 1) / NT), dim3(NT), 0, s, f.pool,
                     pair_target, pair_ld, pair_node, pair_start, contrib, pairs, dminv, vol, lam,
                     mu, R, W, plastic, flags, T, mass, body_reach, body_inertia, pair_body,
                     f.refresh, damping);
}

void factor_zero(const FactorDev &f, const uint64_t *tasks, uint32_t count, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_zero, dim3(count), dim3(NT), 0, s, f, tasks);
}
void factor_extend_add(const FactorDev &f, const uint32_t *parent, const uint64_t *tasks,
                       uint32_t count, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_extend_add, dim3(count), dim3(NT), 0, s, f, parent, tasks);
}
void factor_potrf(const FactorDev &f, const uint32_t *tasks, uint32_t count, uint32_t panel,
                  hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_potrf, dim3(count), dim3(64), 0, s, f, tasks, panel);
}
void factor_trsm(const FactorDev &f, const uint64_t *tasks, uint32_t count, uint32_t panel,
                 hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_trsm, dim3(count), dim3(NT), 0, s, f, tasks, panel);
}
void factor_syrk(const FactorDev &f, const uint64_t *tasks, uint32_t count, uint32_t panel,
                 hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_syrk, dim3(count), dim3(NT), 0, s, f, tasks, panel);
}
void factor_trtri(const FactorDev &f, const uint64_t *tasks, uint32_t count, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_trtri, dim3(count), dim3(NT), 0, s, f, tasks);
}

template <class Real>
void SolveLaunch<Real>::pack(const FactorDev &f, const uint64_t *w21, uint32_t nw21,
                             const uint64_t *m11, uint32_t nm11, G *gr, G *gc,
                             hipStream_t s) {
  if (nw21) hipLaunchKernelGGL(k_pack_w21<Real>, dim3(nw21), dim3(NT), 0, s, f, w21, gr, gc);
  if (nm11) hipLaunchKernelGGL(k_pack_m11<Real>, dim3(nm11), dim3(NT), 0, s, f, m11, gr, gc);
}
template <class Real>
void SolveLaunch<Real>::gather(const FactorDev &f, const uint64_t *tasks, uint32_t count, const Real *rin,
                               const int *done, int parity, uint32_t B, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_gather<Real>, dim3(count), dim3(NT), 0, s, f, tasks, rin, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::xgather(const FactorDev &f, const uint64_t *tasks, uint32_t count, const int *done,
                                int parity, uint32_t B, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_xgather, dim3(count), dim3(NT), 0, s, f, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::forward(const FactorDev &f, const G *gr, const uint64_t *tasks,
                                uint32_t count, const Real *rin, const int *done, int parity, uint32_t B,
                                hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_forward<Real>, dim3(count), dim3(NT), 0, s, f, gr, tasks, rin, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::forward_lanes(const FactorDev &f, const G *gc, const uint64_t *tasks,
                                      uint32_t count, const Real *rin, const int *done, int parity, uint32_t B,
                                      hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_forward_lanes<Real>, dim3(count), dim3(NT), 0, s, f, gc, tasks, rin, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::backward_lanes(const FactorDev &f, const G *gr, const uint64_t *tasks,
                                       uint32_t count, const int *done, int parity, uint32_t B,
                                       hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_backward_lanes<Real>, dim3(count), dim3(NT), 0, s, f, gr, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::forward_split(const FactorDev &f, const G *gc, const uint64_t *tasks,
                                      uint32_t count, const Real *rin, const int *done, int parity, uint32_t B,
                                      hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_forward_split<Real>, dim3(count), dim3(NT), 0, s, f, gc, tasks, rin, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::backward_split(const FactorDev &f, const G *gr, const uint64_t *tasks,
                                       uint32_t count, const int *done, int parity, uint32_t B,
                                       hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_backward_split<Real>, dim3(count), dim3(NT), 0, s, f, gr, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::backward(const FactorDev &f, const G *gc, const uint64_t *tasks,
                                 uint32_t count, const int *done, int parity, uint32_t B,
                                 hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_backward<Real>, dim3(count), dim3(NT), 0, s, f, gc, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::finish(const FactorDev &f, Real *z, uint32_t N, const uint8_t *anchored,
                               const uint32_t *node_body, const int *done, int parity, uint32_t B,
                               const double *frame, hipStream_t s) {
  if (N)
    hipLaunchKernelGGL(k_solve_finish<Real>, dim3((N + NT - 1) / NT), dim3(NT), 0, s, f, z, N, anchored, node_body,
                       done, parity, B, frame);
}

template struct FactorLaunch<float>;
template struct FactorLaunch<double>;
template struct SolveLaunch<float>;
template struct SolveLaunch<double>;

}  // namespace tmi
