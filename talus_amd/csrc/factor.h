// factor.h -- supernodal block Cholesky preconditioner of libtalus_mi.
//
// Replaces the reference's SystemFactor (src/talus/deformable/factor/
// factor.cc:302-314, 318-484, 549-704: Eigen SimplicialLLT with AMD ordering
// of the explicit FP64 system, used as the PCG preconditioner).  Here:
//   ordering   geometric nested dissection of the free nodes (split at the
//              widest coordinate gap near the median of the longest axis,
//              separator = the left nodes touching the right part), leaves
//              of at most `leaf` nodes; every leaf and separator is one
//              dense supernode ("front") in 3x3 node blocks
//   symbolic   rows(s) = cols(s) + later nodes adjacent to cols(s) or in a
//              child's update rows; parent = front of the first update row
//   numeric    multifrontal, level by level from the leaves: assemble the
//              FP64 element blocks h^2 R K R^T + inertia m I, extend-add the
//              children's update matrices, then blocked right-looking
//              Cholesky of each front (POTRF of 64x64 diagonal blocks, TRSM
//              and SYRK/GEMM updates on FP64 MFMA tiles)
//   solve      the diagonal block of each front is inverted explicitly
//              (M11 = L11^-1) and packed with W21 = L21 M11 into
//              G = [M11; W21] (m x k, stored row-major and column-major):
//              the forward pass of a front is one GEMV [y; w] = G v1,
//              u = v2 - w, and the backward pass one transposed GEMV
//              x1 = G^T [y; -x2].  Column j is stored divided by s_j, the
//              power of two nearest M11[j][j] (exact scaling), as FP16 in
//              the mixed path and FP64 in the FP64 path; the solves scale
//              v1 by s on the way in (forward) and x1 by s on the way out
//              (backward), accumulating in FP64
// All sums run in a fixed order: the factor and every solve are
// bit-reproducible.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "host.h"

namespace tmi {

// storage type of the solve matrix G per element precision: the mixed path
// halves the bytes the preconditioner streams (it only shapes the PCG
// iterates; the solve still converges to the requested tolerance)
template <class Real>
struct GStore {
  using type = Real;
};
template <>
struct GStore<float> {
  using type = __half;
};

constexpr int FNB = 64;   // panel / tile width in DOF
constexpr uint32_t DEFAULT_LEAF = 64;  // nodes per leaf front of the nested dissection
constexpr uint32_t ROWS_GROUP = 8;  // rows-shape solve workgroup: 8 rows/columns, 4 wavefronts along the other dimension
constexpr uint32_t GCHUNK = 256;     // front rows per gather workgroup
constexpr uint32_t ZCHUNK = 65536;  // entries zeroed per workgroup before a factorization
constexpr int KLANES = 512;   // fronts with k <= KLANES solve one row/column per lane
constexpr int MLANES = 2048;  // ... backward also needs m <= MLANES (LDS)
constexpr uint32_t LANES_MIN_GROUPS = 2048;  // default plan: lanes when a level has this many
// G's storage per front is padded to multiples of 8 rows and columns (the
// solve kernels read 16-byte tiles of up to 8 entries)
inline uint32_t pad8(uint32_t n) { return (n + 7u) & ~7u; }

struct FrontSym {
  uint32_t ncol = 0;   // nodes eliminated by this front
  uint32_t nrow = 0;   // rows in nodes, the first ncol are the columns
  uint32_t row0 = 0;   // offset into Symbolic::rows
  int32_t parent = -1;
  uint32_t level = 0;  // 0 = leaf
};

// Symbolic analysis of one body (body-local node indices).
struct Symbolic {
  uint32_t nfree = 0;
  std::vector<int32_t> pos;      // node -> elimination index, -1 when anchored
  std::vector<uint32_t> perm;    // elimination index -> node
  std::vector<FrontSym> fronts;  // postorder: children before parents
  std::vector<uint32_t> rows;    // elimination indices of each front's rows
  std::vector<uint32_t> front_of;      // elimination index -> front
  std::vector<uint32_t> child_start;   // CSR of children in ascending order
  std::vector<uint32_t> children;
  uint32_t levels = 0;
  double flops = 0.0;            // dense supernodal factorization flops
  uint64_t factor_entries = 0;   // sum over fronts of k*m - k(k-1)/2 (DOF)
};

Symbolic analyze(const Body &body, uint32_t leaf);

// Device tables for every body of a context (fronts of all bodies are
// concatenated; levels align at the leaves).
struct FactorTables {
  uint32_t fronts = 0, levels = 0, bodies = 0;
  // per front
  std::vector<uint32_t> k, m, body, parent, rows_off, map_off, level;
  std::vector<uint64_t> off, minv_off, vec_off, yoff;
  std::vector<uint32_t> rows;     // global node index of every front row
  std::vector<uint32_t> map;      // per update row: node-row index in the parent
  std::vector<uint32_t> child_start, children;
  // element-pair assembly of the original matrix (lower node blocks)
  std::vector<uint64_t> pair_target;  // FP64 offset of the 3x3 block (col-major, ld m)
  std::vector<uint32_t> pair_ld;      // front m of the target
  std::vector<uint32_t> pair_node;    // global node for the diagonal mass, or ~0u
  std::vector<uint32_t> pair_body;
  std::vector<uint32_t> pair_start;   // CSR into contrib
  std::vector<uint32_t> contrib;      // (global tet << 4) | (a << 2) | b
  // task lists, each a CSR over its launch index
  std::vector<uint32_t> lvl_fronts_start, lvl_fronts;        // fronts per level
  std::vector<uint32_t> ea_start;                            // per (level, round)
  std::vector<uint64_t> ea_tasks;                            // child | ti<<32 | tj<<48
  std::vector<uint32_t> ea_rounds;                           // rounds per level
  std::vector<uint32_t> fac_start;                           // per (level, panel) launch
  std::vector<uint32_t> fac_panels;                          // panels per level
  std::vector<uint32_t> potrf_start, potrf_tasks;            // fronts
  std::vector<uint32_t> trsm_start;
  std::vector<uint64_t> trsm_tasks;                          // front | rowtile<<32
  std::vector<uint32_t> syrk_start;
  std::vector<uint64_t> syrk_tasks;                          // front | I<<32 | J<<48
  std::vector<uint32_t> trtri_start;                         // per row block i
  std::vector<uint64_t> trtri_tasks;                         // front | i<<32 | j<<48
  std::vector<uint32_t> srow_start, scol_start;              // per level
  std::vector<uint64_t> g_off;                               // per front, m*k
  std::vector<uint64_t> w21_tasks, pack_tasks;               // front | I<<32 | J<<48
  std::vector<uint64_t> fwd_tasks, bwd_tasks;                // front | group<<32 | size<<48
  std::vector<uint32_t> fwd_start, bwd_start;
  std::vector<uint64_t> fwl_tasks, bwl_tasks;                // small fronts: front | chunk<<32
  std::vector<uint32_t> fwl_start, bwl_start;
  std::vector<uint64_t> fwr_tasks, bwr_tasks;                // the other fronts, rows per wave
  std::vector<uint32_t> fwr_start, bwr_start;
  std::vector<uint64_t> zero_tasks;  // front | chunk<<32 | array<<48, ZCHUNK entries per chunk
  std::vector<uint64_t> fws_tasks, bws_tasks;                // small fronts, 64 rows/columns per workgroup
  std::vector<uint32_t> fws_start, bws_start;
  // the rows-shape kernels read a front vector gathered once per front
  // (forward: r and the children's updates; backward: x2 at the update
  // rows): front | chunk<<32, GCHUNK rows per chunk, for every front of a
  // level (plan 0) or the fronts too large for the lane shapes (fwr/bwr)
  std::vector<uint64_t> gat_all, gat_big, xg_all, xg_big;
  std::vector<uint32_t> gat_all_start, gat_big_start, xg_all_start, xg_big_start;
  // default solve plan, 2 per level (forward, backward): 0 = rows-per-
  // wavefront kernel throughout; for the small fronts of the level 1 = one
  // row/column per lane, 256 per workgroup; 2 = 64 rows/columns per
  // workgroup with the other dimension split over its 4 wavefronts
  std::vector<uint8_t> plan;
  // pull lists: for every front node row (rows_off indexing), the vec-pool
  // offsets of the child update entries (DOF 0) that add into it
  std::vector<uint32_t> src_start;
  std::vector<uint64_t> src;
  uint64_t pool = 0, minv_pool = 0, vec_pool = 0, y_pool = 0, g_pool = 0;
  double flops = 0.0;
  uint64_t factor_entries = 0;
};

// Builds the tables for bodies whose node/tet ranges start at n0/t0.
void build_tables(const std::vector<Symbolic> &sym, const std::vector<Body> &bodies,
                  const std::vector<uint32_t> &n0, const std::vector<uint32_t> &t0,
                  FactorTables &out);

// Device side.
struct FactorDev {
  uint32_t fronts = 0;
  const uint32_t *k, *m, *body, *rows_off, *map_off, *child_start, *children, *rows, *map;
  const uint32_t *src_start;
  const uint64_t *src;
  const uint64_t *off, *minv_off, *vec_off, *yoff, *g_off;
  double *pool, *minv, *vec, *y, *xsol;
  double *gscale;      // per front column (yoff indexing): s_j of G's column j
  int *failed;         // per body
  const int *refresh;  // per body: nonzero = this factorization rebuilds its fronts
};

// Launch helpers (definitions in factor.hip).  `conn/dminv/...` are the
// engine's element arrays; R is the committed rotation and W the
// neo-Hookean volume frame, both in the element precision; plastic the FP64
// plastic planes (Fp^-1 - I for plastic neo-Hookean tets).
// Stiffness damping of the factor (factor.cc:320-375): the damping blocks
// tau P R K R^T P of the damped tets, P fitted to the positions x at the
// factorization, weighted by each body's damping coefficient; tau null when
// no tet is damped.
template <class Real>
struct FactorDamping {
  const double *x = nullptr;
  const uint4 *conn = nullptr;
  const Real *tau = nullptr;
  const double *body_damp = nullptr;
};

template <class Real>
struct FactorLaunch {
  static void assemble(const FactorDev &f, const uint64_t *pair_target, const uint32_t *pair_ld,
                       const uint32_t *pair_node, const uint32_t *pair_start,
                       const uint32_t *contrib, uint32_t pairs, const double *dminv,
                       const double *vol, const double *lam, const double *mu, const Real *R,
                       const Real *W, const double *plastic, const uint8_t *flags, uint32_t T,
                       const double *mass, const double *body_reach,
                       const double *body_inertia, const uint32_t *pair_body,
                       const FactorDamping<Real> &damping, hipStream_t s);
};

void factor_extend_add(const FactorDev &f, const uint32_t *parent, const uint64_t *tasks,
                       uint32_t count, hipStream_t s);
// zeroes the front storage (pool) and M11 blocks (minv) of the refreshed
// bodies' fronts; tasks = front | chunk<<32 | array<<48 (0 pool, 1 minv)
void factor_zero(const FactorDev &f, const uint64_t *tasks, uint32_t count, hipStream_t s);
void factor_potrf(const FactorDev &f, const uint32_t *tasks, uint32_t count, uint32_t panel,
                  hipStream_t s);
void factor_trsm(const FactorDev &f, const uint64_t *tasks, uint32_t count, uint32_t panel,
                 hipStream_t s);
void factor_syrk(const FactorDev &f, const uint64_t *tasks, uint32_t count, uint32_t panel,
                 hipStream_t s);
void factor_trtri(const FactorDev &f, const uint64_t *tasks, uint32_t count, hipStream_t s);

template <class Real>
struct SolveLaunch {
  using G = typename GStore<Real>::type;
  // G = [M11; L21 M11] in row-major (gr) and column-major (gc) copies
  static void pack(const FactorDev &f, const uint64_t *w21, uint32_t nw21, const uint64_t *m11,
                   uint32_t nm11, G *gr, G *gc, hipStream_t s);
  // forward and backward passes of z = (L L^T)^-1 r, one level per call, for
  // every body whose convergence flag (done[parity*B + body]) is clear;
  // parity < 0 applies to every body.  The forward kernels gather their
  // fronts' right-hand side (r and the children's updates) themselves.
  // front vectors of the rows shape: gather (r and the children's updates)
  // before the forward pass, xgather (x2) before the backward pass
  static void gather(const FactorDev &f, const uint64_t *tasks, uint32_t count, const Real *rin,
                     const int *done, int parity, uint32_t B, hipStream_t s);
  static void xgather(const FactorDev &f, const uint64_t *tasks, uint32_t count, const int *done,
                      int parity, uint32_t B, hipStream_t s);
  static void forward(const FactorDev &f, const G *gr, const uint64_t *tasks, uint32_t count,
                      const Real *rin, const int *done, int parity, uint32_t B, hipStream_t s);
  static void forward_lanes(const FactorDev &f, const G *gc, const uint64_t *tasks,
                            uint32_t count, const Real *rin, const int *done, int parity, uint32_t B,
                            hipStream_t s);
  static void backward(const FactorDev &f, const G *gc, const uint64_t *tasks, uint32_t count,
                       const int *done, int parity, uint32_t B, hipStream_t s);
  static void backward_lanes(const FactorDev &f, const G *gr, const uint64_t *tasks,
                             uint32_t count, const int *done, int parity, uint32_t B,
                             hipStream_t s);
  static void forward_split(const FactorDev &f, const G *gc, const uint64_t *tasks,
                            uint32_t count, const Real *rin, const int *done, int parity, uint32_t B,
                            hipStream_t s);
  static void backward_split(const FactorDev &f, const G *gr, const uint64_t *tasks,
                             uint32_t count, const int *done, int parity, uint32_t B,
                             hipStream_t s);
  // z = F x (F per body, 9 doubles row-major, or the identity when null)
  static void finish(const FactorDev &f, Real *z, uint32_t N, const uint8_t *anchored,
                     const uint32_t *node_body, const int *done, int parity, uint32_t B,
                     const double *frame, hipStream_t s);
};

}  // namespace tmi
