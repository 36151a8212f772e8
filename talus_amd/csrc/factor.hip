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
constexpr int LDP = FNB + 1;

__device__ __forceinline__ uint32_t lo32(uint64_t t) { return uint32_t(t & 0xffffffffu); }
__device__ __forceinline__ uint32_t mid16(uint64_t t) { return uint32_t((t >> 32) & 0xffffu); }
__device__ __forceinline__ uint32_t hi16(uint64_t t) { return uint32_t(t >> 48); }

// ------------------------------------------------------------- assembly

// One lower node block of A = inertia M + h^2 sum_e R K_e R^T per thread,
// summed over its tets in ascending order (factor.cc:32-88 precise blocks).
template <class Real>
__global__ __launch_bounds__(NT) void k_assemble_pairs(
    double *pool, const uint64_t *pair_target, const uint32_t *pair_ld, const uint32_t *pair_node,
    const uint32_t *pair_start, const uint32_t *contrib, uint32_t pairs, const double *dminv,
    const double *vol, const double *lam, const double *mu, const Real *R, uint32_t T,
    const double *mass, const double *reach_of, const double *inertia_of, const uint32_t *pair_body) {
  const uint32_t q = blockIdx.x * NT + threadIdx.x;
  if (q >= pairs) return;
  double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
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
    const double *ga = &g[3 * a], *gb = &g[3 * b];
    const double V = vol[e], l = lam[e], m = mu[e];
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
  }
  const uint32_t body = pair_body[q];
  const double reach = reach_of[body];
#pragma unroll
  for (int k = 0; k < 9; ++k) B[k] *= reach;
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

// --------------------------------------------------------------- extend-add

__global__ __launch_bounds__(NT) void k_extend_add(FactorDev f, const uint32_t *parent,
                                                   const uint64_t *tasks) {
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t c = lo32(t), ti = mid16(t), tj = hi16(t);
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

// Cholesky of the 64x64 diagonal block of panel p in LDS; writes L_pp to
// the front and L_pp^-1 to the (p,p) block of M11.
__global__ __launch_bounds__(NT) void k_potrf(FactorDev f, const uint32_t *tasks, uint32_t panel) {
  __shared__ double a[FNB][LDP];
  __shared__ double x[FNB][LDP];
  const uint32_t fr = tasks[blockIdx.x];
  const uint32_t k = f.k[fr], m = f.m[fr];
  const uint32_t P0 = panel * FNB;
  const int nb = int(min(uint32_t(FNB), k - P0));
  double *A = f.pool + f.off[fr];
  for (int idx = threadIdx.x; idx < FNB * FNB; idx += NT) {
    const int i = idx % FNB, j = idx / FNB;
    a[i][j] = (i < nb && j < nb) ? A[uint64_t(P0 + j) * m + P0 + i] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (threadIdx.x == 0) {
      double d = a[j][j];
      if (!(d > 0.0)) {
        f.failed[f.body[fr]] = 1;
        d = 1.0;
      }
      a[j][j] = sqrt(d);
    }
    __syncthreads();
    const double piv = a[j][j];
    for (int i = j + 1 + threadIdx.x; i < nb; i += NT) a[i][j] /= piv;
    __syncthreads();
    const int w = nb - j - 1;
    for (int idx = threadIdx.x; idx < w * w; idx += NT) {
      const int i = j + 1 + idx / w, l = j + 1 + idx % w;
      if (l <= i) a[i][l] -= a[i][j] * a[l][j];
    }
    __syncthreads();
  }
  // inverse of the lower factor, one column per thread
  if (threadIdx.x < nb) {
    const int c = threadIdx.x;
    for (int i = 0; i < c; ++i) x[i][c] = 0.0;
    x[c][c] = 1.0 / a[c][c];
    for (int i = c + 1; i < nb; ++i) {
      double s = 0.0;
      for (int q = c; q < i; ++q) s += a[i][q] * x[q][c];
      x[i][c] = -s / a[i][i];
    }
  }
  __syncthreads();
  double *M = f.minv + f.minv_off[fr];
  for (int idx = threadIdx.x; idx < nb * nb; idx += NT) {
    const int i = idx % nb, j = idx / nb;
    if (i >= j) {
      A[uint64_t(P0 + j) * m + P0 + i] = a[i][j];
      M[uint64_t(P0 + j) * k + P0 + i] = x[i][j];
    }
  }
}

// --------------------------------------------------------- 64x64 tile GEMM

// acc[4][4] += sum_k As[k][4tr+i] * Bs[k][4tc+j] over kk < kn.
__device__ __forceinline__ void tile_fma(const double (*As)[LDP], const double (*Bs)[LDP], int kn,
                                         double acc[4][4]) {
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  for (int kk = 0; kk < kn; ++kk) {
    double av[4], bv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      av[i] = As[kk][4 * tr + i];
      bv[i] = Bs[kk][4 * tc + i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
  }
}

// TRSM: L_{rt,p} = A_{rt,p} L_pp^-T for the rows below the panel (all m
// rows; tile rt starts 64 rt rows after the panel's last row).
__global__ __launch_bounds__(NT) void k_trsm(FactorDev f, const uint64_t *tasks, uint32_t panel) {
  __shared__ double As[FNB][LDP];  // As[q][r] = A[R0+r][P0+q]
  __shared__ double Xs[FNB][LDP];  // Xs[q][c] = X[c][q] = (L_pp^-1)[c][q]
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), rt = mid16(t);
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
  double acc[4][4] = {};
  tile_fma(As, Xs, nb, acc);
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 4 * tr + i, c = 4 * tc + j;
      if (r < nr && c < nb) A[uint64_t(P0 + c) * m + R0 + r] = acc[i][j];
    }
}

// SYRK/GEMM trailing update A_IJ -= L_Ip L_Jp^T over the rows/columns after
// the panel (tiles counted from the panel's last row).
__global__ __launch_bounds__(NT) void k_syrk(FactorDev f, const uint64_t *tasks, uint32_t panel) {
  __shared__ double Is[FNB][LDP];
  __shared__ double Js[FNB][LDP];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), I = mid16(t), J = hi16(t);
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
  double acc[4][4] = {};
  tile_fma(Is, Js, nb, acc);
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = 4 * tr + i, c = 4 * tc + j;
      if (r < ni && c < nj) A[uint64_t(RJ + c) * m + RI + r] -= acc[i][j];
    }
}

// M11 = L11^-1 by forward substitution on column blocks: for row block i
// and column block j < i, X_ij = -L_ii^-1 sum_{q=j}^{i-1} L_iq X_qj.
__global__ __launch_bounds__(NT) void k_trtri(FactorDev f, const uint64_t *tasks) {
  __shared__ double As[FNB][LDP];  // As[q][r] = L[i-block r][q-block col]
  __shared__ double Bs[FNB][LDP];  // Bs[q][c] = X[q-block row][j-block c]
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), i = mid16(t), j = hi16(t);
  const uint32_t k = f.k[fr], m = f.m[fr];
  const double *L = f.pool + f.off[fr];
  double *X = f.minv + f.minv_off[fr];
  const int ni = int(min(uint32_t(FNB), k - i * FNB)), nj = int(min(uint32_t(FNB), k - j * FNB));
  double acc[4][4] = {};
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
    tile_fma(As, Bs, FNB, acc);
  }
  // X_ij = -Linv_ii * acc
  __syncthreads();
  const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) Bs[4 * tr + a][4 * tc + b] = acc[a][b];  // Bs[r][c] = S[r][c]
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

template <class Real>
__global__ __launch_bounds__(NT) void k_solve_gather(FactorDev f, const uint32_t *fronts,
                                                     const Real *rin, const int *done, int parity,
                                                     uint32_t B) {
  const uint32_t fr = fronts[blockIdx.x];
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  double *v = f.vec + f.vec_off[fr];
  const uint32_t *rows = f.rows + f.rows_off[fr];
  for (uint32_t i = threadIdx.x; i < m; i += NT)
    v[i] = i < k ? double(rin[3 * size_t(rows[i / 3]) + i % 3]) : 0.0;
  for (uint32_t ci = f.child_start[fr]; ci < f.child_start[fr + 1]; ++ci) {
    __syncthreads();
    const uint32_t c = f.children[ci];
    const uint32_t kc = f.k[c], u = f.m[c] - kc;
    const double *vc = f.vec + f.vec_off[c] + kc;
    const uint32_t *map = f.map + f.map_off[c];
    for (uint32_t q = threadIdx.x; q < u; q += NT) v[3 * map[q / 3] + q % 3] += vc[q];
  }
}

// y[rows of tile] = sum_{j<=i} M11[i][j] v[j]; 32 rows x 8 column groups
__global__ __launch_bounds__(NT) void k_solve_m11_rows(FactorDev f, const uint64_t *tasks,
                                                       const int *done, int parity, uint32_t B) {
  __shared__ double part[8][SROW];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), tile = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr];
  const double *M = f.minv + f.minv_off[fr];
  const double *v = f.vec + f.vec_off[fr];
  const uint32_t ri = threadIdx.x % SROW, cg = threadIdx.x / SROW;
  const uint32_t i = tile * SROW + ri;
  double s = 0.0;
  if (i < k)
    for (uint32_t j = cg; j <= i; j += 8) s += M[uint64_t(j) * k + i] * v[j];
  part[cg][ri] = s;
  __syncthreads();
  if (threadIdx.x < SROW && i < k) {
    double y = 0.0;
    for (int g = 0; g < 8; ++g) y += part[g][ri];
    f.y[f.yoff[fr] + i] = y;
  }
}

// v[k+i] -= sum_j L21[i][j] y[j]
__global__ __launch_bounds__(NT) void k_solve_l21_rows(FactorDev f, const uint64_t *tasks,
                                                       const int *done, int parity, uint32_t B) {
  __shared__ double part[8][SROW];
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), tile = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr], m = f.m[fr];
  const double *L = f.pool + f.off[fr];
  const double *y = f.y + f.yoff[fr];
  const uint32_t ri = threadIdx.x % SROW, cg = threadIdx.x / SROW;
  const uint32_t i = tile * SROW + ri;
  double s = 0.0;
  if (k + i < m)
    for (uint32_t j = cg; j < k; j += 8) s += L[uint64_t(j) * m + k + i] * y[j];
  part[cg][ri] = s;
  __syncthreads();
  if (threadIdx.x < SROW && k + i < m) {
    double acc = 0.0;
    for (int g = 0; g < 8; ++g) acc += part[g][ri];
    f.vec[f.vec_off[fr] + k + i] -= acc;
  }
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// y[j] -= sum_i L21[i][j] x[row k+i], one wave per column
__global__ __launch_bounds__(NT) void k_solve_l21_cols(FactorDev f, const uint64_t *tasks,
                                                       const int *done, int parity, uint32_t B) {
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), tile = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr], m = f.m[fr], u = m - k;
  const double *L = f.pool + f.off[fr];
  const uint32_t *rows = f.rows + f.rows_off[fr];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t jj = wave; jj < SROW; jj += NT / 64) {
    const uint32_t j = tile * SROW + jj;
    if (j >= k) break;
    double s = 0.0;
    for (uint32_t i = lane; i < u; i += 64) {
      const uint32_t dof = k + i;
      s += L[uint64_t(j) * m + dof] * f.xsol[3 * size_t(rows[dof / 3]) + dof % 3];
    }
    s = wave_sum64(s);
    if (lane == 0) f.y[f.yoff[fr] + j] -= s;
  }
}

// x[col j] = sum_{i>=j} M11[i][j] z[i], one wave per column
__global__ __launch_bounds__(NT) void k_solve_m11_cols(FactorDev f, const uint64_t *tasks,
                                                       const int *done, int parity, uint32_t B) {
  const uint64_t t = tasks[blockIdx.x];
  const uint32_t fr = lo32(t), tile = mid16(t);
  if (skip(done, parity, B, f.body[fr])) return;
  const uint32_t k = f.k[fr];
  const double *M = f.minv + f.minv_off[fr];
  const double *z = f.y + f.yoff[fr];
  const uint32_t *rows = f.rows + f.rows_off[fr];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t jj = wave; jj < SROW; jj += NT / 64) {
    const uint32_t j = tile * SROW + jj;
    if (j >= k) break;
    double s = 0.0;
    for (uint32_t i = j + lane; i < k; i += 64) s += M[uint64_t(j) * k + i] * z[i];
    s = wave_sum64(s);
    if (lane == 0) f.xsol[3 * size_t(rows[j / 3]) + j % 3] = s;
  }
}

template <class Real>
__global__ __launch_bounds__(NT) void k_solve_finish(FactorDev f, Real *z, uint32_t N,
                                                     const uint8_t *anchored,
                                                     const uint32_t *node_body, const int *done,
                                                     int parity, uint32_t B) {
  const uint32_t n = blockIdx.x * NT + threadIdx.x;
  if (n >= N || skip(done, parity, B, node_body[n])) return;
#pragma unroll
  for (int c = 0; c < 3; ++c)
    z[3 * size_t(n) + c] = anchored[n] ? Real(0) : Real(f.xsol[3 * size_t(n) + c]);
}

}  // namespace

// ---------------------------------------------------------------- launchers

template <class Real>
void FactorLaunch<Real>::assemble(const FactorDev &f, const uint64_t *pair_target,
                                  const uint32_t *pair_ld, const uint32_t *pair_node,
                                  const uint32_t *pair_start, const uint32_t *contrib,
                                  uint32_t pairs, const double *dminv, const double *vol,
                                  const double *lam, const double *mu, const Real *R, uint32_t T,
                                  const double *mass, const double *body_reach,
                                  const double *body_inertia, const uint32_t *pair_body,
                                  hipStream_t s) {
  if (!pairs) return;
  hipLaunchKernelGGL(k_assemble_pairs<Real>, dim3((pairs + NT - 1) / NT), dim3(NT), 0, s, f.pool,
                     pair_target, pair_ld, pair_node, pair_start, contrib, pairs, dminv, vol, lam,
                     mu, R, T, mass, body_reach, body_inertia, pair_body);
}

void factor_extend_add(const FactorDev &f, const uint32_t *parent, const uint64_t *tasks,
                       uint32_t count, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_extend_add, dim3(count), dim3(NT), 0, s, f, parent, tasks);
}
void factor_potrf(const FactorDev &f, const uint32_t *tasks, uint32_t count, uint32_t panel,
                  hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_potrf, dim3(count), dim3(NT), 0, s, f, tasks, panel);
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
void SolveLaunch<Real>::gather(const FactorDev &f, const uint32_t *fronts, uint32_t count,
                               const Real *r, const int *done, int parity, uint32_t B,
                               hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_gather<Real>, dim3(count), dim3(NT), 0, s, f, fronts, r, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::m11_rows(const FactorDev &f, const uint64_t *tasks, uint32_t count,
                                 const int *done, int parity, uint32_t B, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_m11_rows, dim3(count), dim3(NT), 0, s, f, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::l21_rows(const FactorDev &f, const uint64_t *tasks, uint32_t count,
                                 const int *done, int parity, uint32_t B, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_l21_rows, dim3(count), dim3(NT), 0, s, f, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::l21_cols(const FactorDev &f, const uint64_t *tasks, uint32_t count,
                                 const int *done, int parity, uint32_t B, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_l21_cols, dim3(count), dim3(NT), 0, s, f, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::m11_cols(const FactorDev &f, const uint64_t *tasks, uint32_t count,
                                 const int *done, int parity, uint32_t B, hipStream_t s) {
  if (count) hipLaunchKernelGGL(k_solve_m11_cols, dim3(count), dim3(NT), 0, s, f, tasks, done, parity, B);
}
template <class Real>
void SolveLaunch<Real>::finish(const FactorDev &f, Real *z, uint32_t N, const uint8_t *anchored,
                               const uint32_t *node_body, const int *done, int parity, uint32_t B,
                               hipStream_t s) {
  if (N) hipLaunchKernelGGL(k_solve_finish<Real>, dim3((N + NT - 1) / NT), dim3(NT), 0, s, f, z, N, anchored, node_body, done, parity, B);
}

template struct FactorLaunch<float>;
template struct FactorLaunch<double>;
template struct SolveLaunch<float>;
template struct SolveLaunch<double>;

}  // namespace tmi
