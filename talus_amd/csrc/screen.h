// screen.h -- the FP64 neo-Hookean hex terms, the lattice energy ledger and
// the screen solve of a hex lattice on the GPU (screen.hip; not part of the
// C ABI).
//
// Reference (paths relative to the Talus tree):
//   src/talus/dynamics/neohookean/neohookean.cc   cellEnergy, nodeTerms,
//                                                 restStiffness (8 Gauss points)
//   src/talus/dynamics/ledger/ledger.cc:9-52      latticeEnergy
//   src/talus/dynamics/screen/screen.cc           restBodyOf, factorScreen,
//                                                 screenLoad
//
// Layout in HBM (one lattice): node vectors FP64 xyz-interleaved; cells as
// 8 node indices (uint32, cell-major) with per-cell volume, shear and
// shifted modulus (FP64); one slot per (cell, corner) in cell-major order,
// and per node the CSR of its slots in ascending slot order, so a node's
// sum runs in a fixed order (no atomics).  One thread handles one (cell,
// corner): it forms the displacement gradient of its cell at each Gauss
// point from the cell's 8 corners (the 8 threads of a cell read the same
// corners, served from L1) and accumulates its own corner's terms.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "../../include/talus_mi.h"
#include "hexlattice.h"

namespace tmi {

// moduliOf (neohookean.cc:82-90): mu and lambda + mu, nu clamped to [0, 0.49]
void snh_moduli(float youngs, float poisson, double *shear, double *shifted);

// restBodyOf (screen.cc:26-83) of a lattice's rest mesh
void rest_body_of(const LatticeBody &b, talus_mi_rest_body &out);

class HexScreen {
 public:
  explicit HexScreen(hipStream_t stream);
  ~HexScreen();
  HexScreen(const HexScreen &) = delete;
  HexScreen &operator=(const HexScreen &) = delete;

  // uploads the lattice's mesh when it is not the one bound (first call, a
  // tear); the factor-free solve keeps nothing else between calls
  void bind(const LatticeBody &b);
  // nodeTerms of every node (sums over its cells) and cellEnergy of every
  // cell at positions x (3 per node, host); any output may be null
  void terms(const double *x, double *cell_energy, double *gradient, double *hessian);
  // latticeEnergy at positions x and velocities v (host)
  talus_mi_energy energy(const double *x, const double *v, const double *gravity);
  // factorScreen + screenLoad: the rest stiffness plus the screen shifts
  // solved by block-Jacobi PCG in FP64 to `tolerance` relative; returns
  // false when the solve does not reach it within `limit` iterations
  bool screen(const LatticeBody &b, const uint32_t *held, uint32_t nheld, const double *forces,
              double *displacement, double *dent, uint32_t *iterations, double tolerance, uint32_t limit);

 private:
  struct Device;
  hipStream_t stream_;
  Device *dev_ = nullptr;
  uint32_t nodes_ = 0, cells_ = 0;
  std::vector<uint32_t> hex_;  // the bound connectivity
};

}  // namespace tmi
