// screen.hip -- gfx950 kernels and host driver of the FP64 neo-Hookean hex
// terms, the lattice energy ledger and the screen solve (screen.h).
//
// Reference (paths relative to the Talus tree):
//   k_hex_terms     neohookean.cc:92-196 (Gauss gradients, cellEnergy,
//                   nodeTerms) with ledger.cc:29-47's corner-0 shift for the
//                   cell energy
//   k_hex_ledger    ledger.cc:9-27 (kinetic of free nodes, gravity of all)
//   k_screen_*      screen.cc:85-254: the rest stiffness (restStiffness,
//                   neohookean.cc:198-229) applied matrix-free,
//                   r . (K + shift) p with the shifts of factorScreen, and
//                   a block-Jacobi PCG in place of SimplicialLDLT (the same
//                   solution of the same SPD system, to the solve's
//                   tolerance)
#include "screen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "device.h"
#include "tmath.h"

namespace tmi {

namespace {

constexpr int NT = 256;
constexpr double GAUSS = 0.5773502691896258;  // 1 / sqrt(3)
constexpr double SCREEN_SHIFT = 1.0e-9;       // screenShift (screen.cc:21)
constexpr double ANCHOR_SCALE = 1.0e6;        // anchorScale (screen.cc:22)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// fixed-order block sum, valid in every thread
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < NT / 64; ++w) s += lds[w];
  return s;
}

// shape gradient of corner c at Gauss point p (shapeGradients at
// GAUSS * cornerSign(p), neohookean.cc:92-123)
__device__ __forceinline__ void corner_gradient(uint32_t c, uint32_t p, const double *half, double *g) {
  const double s0 = (c & 1u) ? 1.0 : -1.0, s1 = (c & 2u) ? 1.0 : -1.0, s2 = (c & 4u) ? 1.0 : -1.0;
  const double a0 = GAUSS * ((p & 1u) ? 1.0 : -1.0), a1 = GAUSS * ((p & 2u) ? 1.0 : -1.0),
               a2 = GAUSS * ((p & 4u) ? 1.0 : -1.0);
  g[0] = 0.125 * s0 * (1.0 + s1 * a1) * (1.0 + s2 * a2) / half[0];
  g[1] = 0.125 * s1 * (1.0 + s0 * a0) * (1.0 + s2 * a2) / half[1];
  g[2] = 0.125 * s2 * (1.0 + s0 * a0) * (1.0 + s1 * a1) / half[2];
}

struct Cells {
  const uint32_t *hex;
  const double *rest, *vol, *shear, *shifted;
  double half[3];
  uint32_t cells;
};

// displacement gradient H = sum_o u_o g_o(p)^T of a cell at Gauss point p
__device__ __forceinline__ void offset_at(const double *u, uint32_t p, const double *half, double *H) {
#pragma unroll
  for (int k = 0; k < 9; ++k) H[k] = 0.0;
#pragma unroll
  for (uint32_t o = 0; o < 8; ++o) {
    double g[3];
    corner_gradient(o, p, half, g);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int k = 0; k < 3; ++k) H[3 * r + k] += u[3 * o + r] * g[k];
  }
}

__device__ __forceinline__ double change_of(const double *H) {
  const double tr = H[0] + H[4] + H[8];
  double hht = 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) hht += H[3 * r + k] * H[3 * k + r];
  return tr + 0.5 * (tr * tr - hht) + det3(H);
}

// One (cell, corner) per thread: the corner's gradient (3) and Hessian
// block (6, upper triangle xx xy xz yy yz zz) into its slot; the corner-0
// thread also writes the cell's energy with corner 0's displacement taken
// out of every corner (latticeEnergy).
__global__ __launch_bounds__(NT) void k_hex_terms(Cells c, const double *x, double *slot, double *cell_energy) {
  const uint32_t t = blockIdx.x * NT + threadIdx.x;
  if (t >= 8u * c.cells) return;
  const uint32_t cell = t >> 3, corner = t & 7u;
  double u[24];
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const uint32_t n = c.hex[8 * size_t(cell) + o];
#pragma unroll
    for (int r = 0; r < 3; ++r) u[3 * o + r] = x[3 * size_t(n) + r] - c.rest[3 * size_t(n) + r];
  }
  const double w = 0.125 * c.vol[cell], mu = c.shear[cell], s = c.shifted[cell];
  double grad[3] = {0.0, 0.0, 0.0}, hess[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t p = 0; p < 8; ++p) {
    double H[9], F[9], C[9], g[3];
    offset_at(u, p, c.half, H);
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
    cof3(F, C);
    const double change = change_of(H);
    corner_gradient(corner, p, c.half, g);
    double b[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double sg = 0.0;
      b[r] = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sg += (mu * (F[3 * r + k] - C[3 * r + k]) + s * change * C[3 * r + k]) * g[k];
        b[r] += C[3 * r + k] * g[k];
      }
      grad[r] += w * sg;
    }
    const double gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    hess[0] += w * (mu * gg + s * b[0] * b[0]);
    hess[1] += w * (s * b[0] * b[1]);
    hess[2] += w * (s * b[0] * b[2]);
    hess[3] += w * (mu * gg + s * b[1] * b[1]);
    hess[4] += w * (s * b[1] * b[2]);
    hess[5] += w * (mu * gg + s * b[2] * b[2]);
  }
  double *o = slot + 9 * size_t(t);
#pragma unroll
  for (int k = 0; k < 3; ++k) o[k] = grad[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) o[3 + k] = hess[k];
  if (corner == 0u) {
    double v[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) v[k] = u[k] - u[k % 3];
    double e = 0.0;
    for (uint32_t p = 0; p < 8; ++p) {
      double H[9];
      offset_at(v, p, c.half, H);
      const double change = change_of(H);
      double hh = 0.0;
#pragma unroll
      for (int k = 0; k < 9; ++k) hh += H[k] * H[k];
      const double tr = H[0] + H[4] + H[8];
      e += w * (0.5 * mu * hh - mu * (change - tr) + 0.5 * s * change * change);
    }
    cell_energy[cell] = e;
  }
}

// per node: the sums of its slots' gradients and Hessian blocks in slot order
__global__ __launch_bounds__(NT) void k_hex_gather(uint32_t nodes, const uint32_t *start, const uint32_t *slots,
                                                   const double *slot, double *gradient, double *hessian) {
  const uint32_t n = blockIdx.x * NT + threadIdx.x;
  if (n >= nodes) return;
  double g[3] = {0.0, 0.0, 0.0}, h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t k = start[n]; k < start[n + 1]; ++k) {
    const double *o = slot + 9 * size_t(slots[k]);
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i] += o[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) h[i] += o[3 + i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) gradient[3 * size_t(n) + i] = g[i];
  const double full[9] = {h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]};
#pragma unroll
  for (int i = 0; i < 9; ++i) hessian[9 * size_t(n) + i] = full[i];
}

// per block: kinetic energy of the free nodes, -m g.x of every node and the
// cells' energies (3 partials per block)
__global__ __launch_bounds__(NT) void k_hex_ledger(uint32_t nodes, uint32_t cells, const double *mass,
                                                   const uint8_t *anchored, const double *x, const double *v,
                                                   double g0, double g1, double g2, const double *cell_energy,
                                                   double *partial) {
  __shared__ double lds[NT / 64];
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  double ke = 0.0, pe = 0.0, el = 0.0;
  if (i < nodes) {
    const double m = mass[i];
    const double *vi = v + 3 * size_t(i), *xi = x + 3 * size_t(i);
    if (!anchored[i]) ke = 0.5 * m * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    pe = -m * (g0 * xi[0] + g1 * xi[1] + g2 * xi[2]);
  }
  if (i < cells) el = cell_energy[i];
  ke = block_sum(ke, lds);
  pe = block_sum(pe, lds);
  el = block_sum(el, lds);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x] = ke;
    partial[3 * blockIdx.x + 1] = pe;
    partial[3 * blockIdx.x + 2] = el;
  }
}

// ---------------------------------------------------------------- screen

// One (cell, corner) per thread: y_c = sum_o K[c][o] p_o of the rest
// stiffness, = w sum_p (mu (D + D^T) a + (shifted - mu) tr(D) a) with
// D = sum_o p_o g_o(p)^T and a = g_c(p)
__global__ __launch_bounds__(NT) void k_screen_cells(Cells c, const double *p, double *slot) {
  const uint32_t t = blockIdx.x * NT + threadIdx.x;
  if (t >= 8u * c.cells) return;
  const uint32_t cell = t >> 3, corner = t & 7u;
  double u[24];
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const uint32_t n = c.hex[8 * size_t(cell) + o];
#pragma unroll
    for (int r = 0; r < 3; ++r) u[3 * o + r] = p[3 * size_t(n) + r];
  }
  const double w = 0.125 * c.vol[cell], mu = c.shear[cell], lam = c.shifted[cell] - c.shear[cell];
  double y[3] = {0.0, 0.0, 0.0};
  for (uint32_t q = 0; q < 8; ++q) {
    double D[9], a[3];
    offset_at(u, q, c.half, D);
    corner_gradient(corner, q, c.half, a);
    const double tr = D[0] + D[4] + D[8];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double v = lam * tr * a[r];
#pragma unroll
      for (int k = 0; k < 3; ++k) v += mu * (D[3 * r + k] + D[3 * k + r]) * a[k];
      y[r] += w * v;
    }
  }
  double *o = slot + 3 * size_t(t);
#pragma unroll
  for (int r = 0; r < 3; ++r) o[r] = y[r];
}

// the diagonal block K[c][c] = w sum_p (mu |a|^2 I + shifted a a^T) of each
// (cell, corner), upper triangle
__global__ __launch_bounds__(NT) void k_screen_diag(Cells c, double *slot) {
  const uint32_t t = blockIdx.x * NT + threadIdx.x;
  if (t >= 8u * c.cells) return;
  const uint32_t cell = t >> 3, corner = t & 7u;
  const double w = 0.125 * c.vol[cell], mu = c.shear[cell], s = c.shifted[cell];
  double h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t q = 0; q < 8; ++q) {
    double a[3];
    corner_gradient(corner, q, c.half, a);
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    h[0] += w * (mu * aa + s * a[0] * a[0]);
    h[1] += w * (s * a[0] * a[1]);
    h[2] += w * (s * a[0] * a[2]);
    h[3] += w * (mu * aa + s * a[1] * a[1]);
    h[4] += w * (s * a[1] * a[2]);
    h[5] += w * (mu * aa + s * a[2] * a[2]);
  }
  double *o = slot + 6 * size_t(t);
#pragma unroll
  for (int k = 0; k < 6; ++k) o[k] = h[k];
}

__global__ __launch_bounds__(NT) void k_screen_diag_gather(uint32_t nodes, const uint32_t *start,
                                                           const uint32_t *slots, const double *slot,
                                                           double *block) {
  const uint32_t n = blockIdx.x * NT + threadIdx.x;
  if (n >= nodes) return;
  double h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t k = start[n]; k < start[n + 1]; ++k)
#pragma unroll
    for (int i = 0; i < 6; ++i) h[i] += slot[6 * size_t(slots[k]) + i];
#pragma unroll
  for (int i = 0; i < 6; ++i) block[6 * size_t(n) + i] = h[i];
}

// q = sum of the node's slots + shift p; partial p.q
__global__ __launch_bounds__(NT) void k_screen_gather(uint32_t nodes, const uint32_t *start, const uint32_t *slots,
                                                      const double *slot, const double *shift, const double *p,
                                                      double *q, double *partial) {
  __shared__ double lds[NT / 64];
  const uint32_t n = blockIdx.x * NT + threadIdx.x;
  double pq = 0.0;
  if (n < nodes) {
    double y[3] = {0.0, 0.0, 0.0};
    for (uint32_t k = start[n]; k < start[n + 1]; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i) y[i] += slot[3 * size_t(slots[k]) + i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const size_t j = 3 * size_t(n) + i;
      y[i] += shift[n] * p[j];
      q[j] = y[i];
      pq += p[j] * y[i];
    }
  }
  pq = block_sum(pq, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = pq;
}

__device__ __forceinline__ void apply_block(const double *P, const double *r, double *z) {
#pragma unroll
  for (int i = 0; i < 3; ++i) z[i] = P[3 * i] * r[0] + P[3 * i + 1] * r[1] + P[3 * i + 2] * r[2];
}

// start: x = 0, r = b, z = P r, p = z; partials r.z and b.b
__global__ __launch_bounds__(NT) void k_screen_init(uint32_t nodes, const double *pinv, const double *b, double *x,
                                                    double *r, double *z, double *p, double *partial) {
  __shared__ double lds[NT / 64];
  const uint32_t n = blockIdx.x * NT + threadIdx.x;
  double rz = 0.0, bb = 0.0;
  if (n < nodes) {
    const size_t j = 3 * size_t(n);
    double zz[3];
    apply_block(pinv + 9 * size_t(n), b + j, zz);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[j + i] = 0.0;
      r[j + i] = b[j + i];
      z[j + i] = zz[i];
      p[j + i] = zz[i];
      rz += b[j + i] * zz[i];
      bb += b[j + i] * b[j + i];
    }
  }
  rz = block_sum(rz, lds);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = rz;
    partial[2 * blockIdx.x + 1] = bb;
  }
}

// x += alpha p, r -= alpha q, z = P r; partials r.z and r.r
__global__ __launch_bounds__(NT) void k_screen_update(uint32_t nodes, double alpha, const double *pinv,
                                                      const double *p, const double *q, double *x, double *r,
                                                      double *z, double *partial) {
  __shared__ double lds[NT / 64];
  const uint32_t n = blockIdx.x * NT + threadIdx.x;
  double rz = 0.0, rr = 0.0;
  if (n < nodes) {
    const size_t j = 3 * size_t(n);
    double rn[3], zz[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      x[j + i] += alpha * p[j + i];
      rn[i] = r[j + i] - alpha * q[j + i];
      r[j + i] = rn[i];
    }
    apply_block(pinv + 9 * size_t(n), rn, zz);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      z[j + i] = zz[i];
      rz += rn[i] * zz[i];
      rr += rn[i] * rn[i];
    }
  }
  rz = block_sum(rz, lds);
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = rz;
    partial[2 * blockIdx.x + 1] = rr;
  }
}

__global__ __launch_bounds__(NT) void k_screen_direction(uint32_t nodes, double beta, const double *z, double *p) {
  const uint32_t i = blockIdx.x * NT + threadIdx.x;
  if (i >= 3 * nodes) return;
  p[i] = z[i] + beta * p[i];
}

inline unsigned blocks(size_t n) { return unsigned((n + NT - 1) / NT); }

template <class T>
T *device_copy(const std::vector<T> &h) {
  T *p = nullptr;
  HIPCHECK(hipMalloc(&p, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) HIPCHECK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

template <class T>
T *device_alloc(size_t n) {
  T *p = nullptr;
  HIPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return p;
}

bool inverse3(const double *a, double *o) {
  double c[9];
  cof3(a, c);
  const double d = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
  if (!(d != 0.0) || !std::isfinite(d)) return false;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) o[3 * r + k] = c[3 * k + r] / d;
  return true;
}

}  // namespace

void snh_moduli(float youngs, float poisson, double *shear, double *shifted) {
  const double nu = std::clamp(double(poisson), 0.0, 0.49);
  const double E = double(youngs);
  const double first = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  *shear = E / (2.0 * (1.0 + nu));
  *shifted = first + *shear;
}

void rest_body_of(const LatticeBody &b, talus_mi_rest_body &out) {
  std::memset(&out, 0, sizeof out);
  double low[3], high[3];
  for (int k = 0; k < 3; ++k) {
    low[k] = std::numeric_limits<double>::infinity();
    high[k] = -std::numeric_limits<double>::infinity();
  }
  for (uint32_t n = 0; n < b.nodes; ++n) {
    const double m = double(b.mass[n]);
    out.total += m;
    for (int k = 0; k < 3; ++k) {
      const double p = b.rest[3 * size_t(n) + k];
      out.centre[k] += m * p;
      low[k] = std::min(low[k], p);
      high[k] = std::max(high[k], p);
    }
  }
  for (int k = 0; k < 3; ++k) out.centre[k] /= out.total;
  for (uint32_t n = 0; n < b.nodes; ++n) {
    const double m = double(b.mass[n]);
    double a[3];
    for (int k = 0; k < 3; ++k) a[k] = b.rest[3 * size_t(n) + k] - out.centre[k];
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) out.inertia[3 * r + k] += m * ((r == k ? aa : 0.0) - a[r] * a[k]);
  }
  // principal axes: eigenvalues ascending, eigenvectors as columns, a
  // reflection turning the last axis (SelfAdjointEigenSolver, screen.cc:55-66)
  double w[3], V[9];
  eigen_sym3(out.inertia, w, V);
  int order[3] = {0, 1, 2};
  std::stable_sort(order, order + 3, [&](int i, int j) { return w[i] < w[j]; });
  for (int c = 0; c < 3; ++c) {
    out.moments[c] = w[order[c]];
    for (int r = 0; r < 3; ++r) out.axes[3 * r + c] = V[3 * r + order[c]];
  }
  if (det3(out.axes) < 0.0)
    for (int r = 0; r < 3; ++r) out.axes[3 * r + 2] = -out.axes[3 * r + 2];
  out.thickness = std::min({high[0] - low[0], high[1] - low[1], high[2] - low[2]});
}

struct HexScreen::Device {
  Cells cells{};
  uint32_t *hex = nullptr, *inc_start = nullptr, *inc_slot = nullptr;
  double *rest = nullptr, *vol = nullptr, *shear = nullptr, *shifted = nullptr, *mass = nullptr;
  uint8_t *anchored = nullptr;
  double *x = nullptr, *v = nullptr, *slot = nullptr, *cell_e = nullptr, *node_g = nullptr, *node_h = nullptr;
  double *partial = nullptr;
  // PCG vectors, the shifts and the block-Jacobi inverses
  double *b = nullptr, *px = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *shift = nullptr, *pinv = nullptr, *block = nullptr;
  std::vector<void *> owned;

  template <class T>
  T *keep(T *ptr) {
    owned.push_back(ptr);
    return ptr;
  }
  ~Device() {
    for (void *ptr : owned) (void)hipFree(ptr);
  }
};

HexScreen::HexScreen(hipStream_t stream) : stream_(stream) {}

HexScreen::~HexScreen() { delete dev_; }

void HexScreen::bind(const LatticeBody &b) {
  if (dev_ && nodes_ == b.nodes && cells_ == b.hexes && hex_ == b.hex) return;
  delete dev_;
  dev_ = nullptr;
  const uint32_t N = b.nodes, H = b.hexes;
  std::vector<double> vol(H), shear(H), shifted(H), mass(N);
  for (uint32_t c = 0; c < H; ++c) {
    vol[c] = double(b.volume[c]);
    const talus_mi_material &m = b.material_of(c);
    snh_moduli(m.youngs_modulus, m.poissons_ratio, &shear[c], &shifted[c]);
  }
  for (uint32_t n = 0; n < N; ++n) mass[n] = double(b.mass[n]);
  // node -> (cell, corner) slots in ascending slot order
  std::vector<uint32_t> start(N + 1, 0), slots(8 * size_t(H));
  for (size_t s = 0; s < 8 * size_t(H); ++s) ++start[b.hex[s] + 1];
  for (uint32_t n = 0; n < N; ++n) start[n + 1] += start[n];
  {
    std::vector<uint32_t> cursor(start.begin(), start.end() - 1);
    for (size_t s = 0; s < 8 * size_t(H); ++s) slots[cursor[b.hex[s]]++] = uint32_t(s);
  }
  std::vector<uint8_t> anchored(b.anchored.begin(), b.anchored.end());
  anchored.resize(N, 0);
  Device *d = new Device();
  try {
    d->hex = d->keep(device_copy(b.hex));
    d->rest = d->keep(device_copy(b.rest));
    d->vol = d->keep(device_copy(vol));
    d->shear = d->keep(device_copy(shear));
    d->shifted = d->keep(device_copy(shifted));
    d->mass = d->keep(device_copy(mass));
    d->anchored = d->keep(device_copy(anchored));
    d->inc_start = d->keep(device_copy(start));
    d->inc_slot = d->keep(device_copy(slots));
    const size_t n3 = 3 * size_t(N);
    d->x = d->keep(device_alloc<double>(n3));
    d->v = d->keep(device_alloc<double>(n3));
    d->slot = d->keep(device_alloc<double>(9 * 8 * size_t(H)));
    d->cell_e = d->keep(device_alloc<double>(H));
    d->node_g = d->keep(device_alloc<double>(n3));
    d->node_h = d->keep(device_alloc<double>(9 * size_t(N)));
    d->partial = d->keep(device_alloc<double>(3 * size_t(blocks(std::max<size_t>(std::max(N, H), 1)))));
    d->b = d->keep(device_alloc<double>(n3));
    d->px = d->keep(device_alloc<double>(n3));
    d->r = d->keep(device_alloc<double>(n3));
    d->z = d->keep(device_alloc<double>(n3));
    d->p = d->keep(device_alloc<double>(n3));
    d->q = d->keep(device_alloc<double>(n3));
    d->shift = d->keep(device_alloc<double>(N));
    d->pinv = d->keep(device_alloc<double>(9 * size_t(N)));
    d->block = d->keep(device_alloc<double>(6 * size_t(N)));
  } catch (...) {
    delete d;
    throw;
  }
  d->cells.hex = d->hex;
  d->cells.rest = d->rest;
  d->cells.vol = d->vol;
  d->cells.shear = d->shear;
  d->cells.shifted = d->shifted;
  d->cells.cells = H;
  for (int k = 0; k < 3; ++k) d->cells.half[k] = 0.5 * double(b.voxel[k]);
  dev_ = d;
  nodes_ = N;
  cells_ = H;
  hex_ = b.hex;
}

void HexScreen::terms(const double *x, double *cell_energy, double *gradient, double *hessian) {
  Device &d = *dev_;
  const size_t n3 = 3 * size_t(nodes_);
  HIPCHECK(hipMemcpyAsync(d.x, x, n3 * sizeof(double), hipMemcpyHostToDevice, stream_));
  if (cells_)
    hipLaunchKernelGGL(k_hex_terms, dim3(blocks(8 * size_t(cells_))), dim3(NT), 0, stream_, d.cells, d.x, d.slot,
                       d.cell_e);
  if (nodes_)
    hipLaunchKernelGGL(k_hex_gather, dim3(blocks(nodes_)), dim3(NT), 0, stream_, nodes_, d.inc_start, d.inc_slot,
                       d.slot, d.node_g, d.node_h);
  HIPCHECK(hipGetLastError());
  if (cell_energy)
    HIPCHECK(hipMemcpyAsync(cell_energy, d.cell_e, cells_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
  if (gradient) HIPCHECK(hipMemcpyAsync(gradient, d.node_g, n3 * sizeof(double), hipMemcpyDeviceToHost, stream_));
  if (hessian)
    HIPCHECK(hipMemcpyAsync(hessian, d.node_h, 9 * size_t(nodes_) * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHECK(hipStreamSynchronize(stream_));
}

talus_mi_energy HexScreen::energy(const double *x, const double *v, const double *gravity) {
  Device &d = *dev_;
  const size_t n3 = 3 * size_t(nodes_);
  HIPCHECK(hipMemcpyAsync(d.x, x, n3 * sizeof(double), hipMemcpyHostToDevice, stream_));
  HIPCHECK(hipMemcpyAsync(d.v, v, n3 * sizeof(double), hipMemcpyHostToDevice, stream_));
  if (cells_)
    hipLaunchKernelGGL(k_hex_terms, dim3(blocks(8 * size_t(cells_))), dim3(NT), 0, stream_, d.cells, d.x, d.slot,
                       d.cell_e);
  const size_t count = std::max<size_t>(std::max(nodes_, cells_), 1);
  const unsigned nb = blocks(count);
  hipLaunchKernelGGL(k_hex_ledger, dim3(nb), dim3(NT), 0, stream_, nodes_, cells_, d.mass, d.anchored, d.x, d.v,
                     gravity[0], gravity[1], gravity[2], d.cell_e, d.partial);
  HIPCHECK(hipGetLastError());
  std::vector<double> part(3 * size_t(nb));
  HIPCHECK(hipMemcpyAsync(part.data(), d.partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHECK(hipStreamSynchronize(stream_));
  talus_mi_energy e{0.0, 0.0, 0.0};
  for (unsigned i = 0; i < nb; ++i) {
    e.kinetic += part[3 * i];
    e.gravity += part[3 * i + 1];
    e.elastic += part[3 * i + 2];
  }
  return e;
}

bool HexScreen::screen(const LatticeBody &lb, const uint32_t *held, uint32_t nheld, const double *forces,
                       double *displacement, double *dent, uint32_t *iterations, double tolerance,
                       uint32_t limit) {
  Device &d = *dev_;
  const uint32_t N = nodes_;
  const size_t n3 = 3 * size_t(N);
  // the diagonal blocks of the rest stiffness, for the shifts and the
  // block-Jacobi preconditioner
  std::vector<double> block(6 * size_t(N));
  if (cells_)
    hipLaunchKernelGGL(k_screen_diag, dim3(blocks(8 * size_t(cells_))), dim3(NT), 0, stream_, d.cells, d.slot);
  if (N)
    hipLaunchKernelGGL(k_screen_diag_gather, dim3(blocks(N)), dim3(NT), 0, stream_, N, d.inc_start, d.inc_slot,
                       d.slot, d.block);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(block.data(), d.block, block.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHECK(hipStreamSynchronize(stream_));
  // factorScreen's shifts (screen.cc:128-165)
  double largest = -std::numeric_limits<double>::infinity(), heaviest = 0.0;
  for (uint32_t n = 0; n < N; ++n) {
    largest = std::max({largest, block[6 * size_t(n)], block[6 * size_t(n) + 3], block[6 * size_t(n) + 5]});
    heaviest = std::max(heaviest, double(lb.mass[n]));
  }
  std::vector<uint8_t> hold(N, 0);
  for (uint32_t n = 0; n < N; ++n) hold[n] = n < lb.anchored.size() && lb.anchored[n] != 0;
  for (uint32_t k = 0; k < nheld; ++k) hold[held[k]] = 1;
  bool anchored = false;
  std::vector<double> shift(N), pinv(9 * size_t(N));
  for (uint32_t n = 0; n < N; ++n) {
    anchored = anchored || hold[n];
    shift[n] = hold[n] ? ANCHOR_SCALE * largest : SCREEN_SHIFT * largest / heaviest * double(lb.mass[n]);
    const double *h = &block[6 * size_t(n)];
    const double A[9] = {h[0] + shift[n], h[1], h[2], h[1], h[3] + shift[n], h[4], h[2], h[4], h[5] + shift[n]};
    if (!inverse3(A, &pinv[9 * size_t(n)])) {
      std::memset(&pinv[9 * size_t(n)], 0, 9 * sizeof(double));
      pinv[9 * size_t(n)] = pinv[9 * size_t(n) + 4] = pinv[9 * size_t(n) + 8] = 1.0;
    }
  }
  // screenLoad's load (screen.cc:167-200): relieved of the rigid
  // acceleration when nothing is held, zero on held nodes
  talus_mi_rest_body rb;
  rest_body_of(lb, rb);
  double Iinv[9];
  if (!inverse3(rb.inertia, Iinv)) std::memset(Iinv, 0, sizeof Iinv);
  double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0}, linear[3] = {0, 0, 0}, angular[3] = {0, 0, 0};
  for (uint32_t n = 0; n < N; ++n) {
    const double *f = &forces[3 * size_t(n)];
    double a[3];
    for (int k = 0; k < 3; ++k) a[k] = lb.rest[3 * size_t(n) + k] - rb.centre[k];
    for (int k = 0; k < 3; ++k) force[k] += f[k];
    torque[0] += a[1] * f[2] - a[2] * f[1];
    torque[1] += a[2] * f[0] - a[0] * f[2];
    torque[2] += a[0] * f[1] - a[1] * f[0];
  }
  if (!anchored) {
    for (int k = 0; k < 3; ++k) linear[k] = force[k] / rb.total;
    for (int r = 0; r < 3; ++r) angular[r] = Iinv[3 * r] * torque[0] + Iinv[3 * r + 1] * torque[1] + Iinv[3 * r + 2] * torque[2];
  }
  std::vector<double> load(n3, 0.0);
  for (uint32_t n = 0; n < N; ++n) {
    if (hold[n]) continue;
    double a[3];
    for (int k = 0; k < 3; ++k) a[k] = lb.rest[3 * size_t(n) + k] - rb.centre[k];
    const double wa[3] = {angular[1] * a[2] - angular[2] * a[1], angular[2] * a[0] - angular[0] * a[2],
                          angular[0] * a[1] - angular[1] * a[0]};
    const double m = double(lb.mass[n]);
    for (int k = 0; k < 3; ++k) load[3 * size_t(n) + k] = forces[3 * size_t(n) + k] - m * (linear[k] + wa[k]);
  }
  HIPCHECK(hipMemcpyAsync(d.shift, shift.data(), N * sizeof(double), hipMemcpyHostToDevice, stream_));
  HIPCHECK(hipMemcpyAsync(d.pinv, pinv.data(), pinv.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
  HIPCHECK(hipMemcpyAsync(d.b, load.data(), n3 * sizeof(double), hipMemcpyHostToDevice, stream_));
  // block-Jacobi PCG on (K + shift) u = load
  const unsigned nb = blocks(std::max<uint32_t>(N, 1));
  std::vector<double> part(2 * size_t(nb));
  auto sums = [&](double &s0, double &s1) {
    HIPCHECK(hipMemcpyAsync(part.data(), d.partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    s0 = s1 = 0.0;
    for (unsigned i = 0; i < nb; ++i) {
      s0 += part[2 * i];
      s1 += part[2 * i + 1];
    }
  };
  hipLaunchKernelGGL(k_screen_init, dim3(nb), dim3(NT), 0, stream_, N, d.pinv, d.b, d.px, d.r, d.z, d.p, d.partial);
  HIPCHECK(hipGetLastError());
  double rz = 0.0, bb = 0.0;
  sums(rz, bb);
  const double target = tolerance * std::sqrt(bb);
  bool converged = !(bb > 0.0);
  uint32_t it = 0;
  while (!converged && it < limit) {
    if (cells_)
      hipLaunchKernelGGL(k_screen_cells, dim3(blocks(8 * size_t(cells_))), dim3(NT), 0, stream_, d.cells, d.p,
                         d.slot);
    hipLaunchKernelGGL(k_screen_gather, dim3(nb), dim3(NT), 0, stream_, N, d.inc_start, d.inc_slot, d.slot, d.shift,
                       d.p, d.q, d.partial);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(part.data(), d.partial, nb * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
    double pq = 0.0;
    for (unsigned i = 0; i < nb; ++i) pq += part[i];
    if (!(pq > 0.0)) break;
    const double alpha = rz / pq;
    hipLaunchKernelGGL(k_screen_update, dim3(nb), dim3(NT), 0, stream_, N, alpha, d.pinv, d.p, d.q, d.px, d.r, d.z,
                       d.partial);
    HIPCHECK(hipGetLastError());
    double rz_next = 0.0, rr = 0.0;
    sums(rz_next, rr);
    ++it;
    if (std::sqrt(rr) <= target) {
      converged = true;
      break;
    }
    const double beta = rz_next / rz;
    rz = rz_next;
    hipLaunchKernelGGL(k_screen_direction, dim3(blocks(n3)), dim3(NT), 0, stream_, N, beta, d.z, d.p);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipMemcpyAsync(displacement, d.px, n3 * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHECK(hipStreamSynchronize(stream_));
  if (iterations) *iterations = it;
  // the mass-weighted rigid part removed when nothing is held
  // (screen.cc:214-244), dent = max |u| / thickness
  if (!anchored) {
    double sh[3] = {0, 0, 0}, moment[3] = {0, 0, 0}, turn[3];
    for (uint32_t n = 0; n < N; ++n)
      for (int k = 0; k < 3; ++k) sh[k] += double(lb.mass[n]) * displacement[3 * size_t(n) + k];
    for (int k = 0; k < 3; ++k) sh[k] /= rb.total;
    for (uint32_t n = 0; n < N; ++n) {
      double a[3], u[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = lb.rest[3 * size_t(n) + k] - rb.centre[k];
        u[k] = displacement[3 * size_t(n) + k] - sh[k];
      }
      const double m = double(lb.mass[n]);
      moment[0] += m * (a[1] * u[2] - a[2] * u[1]);
      moment[1] += m * (a[2] * u[0] - a[0] * u[2]);
      moment[2] += m * (a[0] * u[1] - a[1] * u[0]);
    }
    for (int r = 0; r < 3; ++r) turn[r] = Iinv[3 * r] * moment[0] + Iinv[3 * r + 1] * moment[1] + Iinv[3 * r + 2] * moment[2];
    for (uint32_t n = 0; n < N; ++n) {
      double a[3];
      for (int k = 0; k < 3; ++k) a[k] = lb.rest[3 * size_t(n) + k] - rb.centre[k];
      const double ta[3] = {turn[1] * a[2] - turn[2] * a[1], turn[2] * a[0] - turn[0] * a[2],
                            turn[0] * a[1] - turn[1] * a[0]};
      for (int k = 0; k < 3; ++k) displacement[3 * size_t(n) + k] -= sh[k] + ta[k];
    }
  }
  double most = 0.0;
  for (uint32_t n = 0; n < N; ++n) {
    const double *u = &displacement[3 * size_t(n)];
    most = std::max(most, std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
  }
  *dent = most / rb.thickness;
  return converged;
}

}  // namespace tmi
