// hexlattice.h -- strainstep hex lattices in libtalus_mi (not part of the
// C ABI): the host element build (hex_build.cpp) and the device-side set of
// lattice bodies stepped by the HIP kernels (lattice.hip).
//
// Reference: src/talus/dynamics/strainstep/element.cc (the 18-row element),
// strainstep.cc:56-126, 551-842 (the substep loop), bodyset.cc:13-86
// (LatticeBody).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "../../include/talus_mi.h"

namespace tmi {

constexpr uint32_t HEX_ROWS = 18;
constexpr uint32_t HEX_INVERSE = HEX_ROWS * HEX_ROWS;

// makeElement's outputs plus the state that carries over between steps.
struct HexElement {
  uint32_t cell = 0;
  uint32_t nodes[8] = {};
  double shape_inverse[9] = {};  // rest shape^-1, row-major (gradients: inv^T s / 8)
  double gradient[24] = {};
  double inv_mass[8] = {};
  double rest_hg[12] = {};
  double bend = 0, size = 0;
  double stiffness[HEX_ROWS] = {}, rate[HEX_ROWS] = {}, softness[HEX_ROWS] = {},
         cap[HEX_ROWS] = {};
  std::array<double, HEX_INVERSE> inverse{};
  double vol_stiffness = 0, vol_rate = 0, vol_softness = 0, vol_cap = 0;
  bool volumetric = false;
  // state
  double impulse[HEX_ROWS] = {};
  double vol_impulse = 0;
  double quat[4] = {1, 0, 0, 0};  // w, x, y, z
};

// ElementDamping (element.h:33-38): the lane's target ratio, the lattice's
// lowest angular frequency, the implicit weight theta and theta h
struct HexDamping {
  double target = 0.0, lowest = 0.0, theta = 1.0, own = 0.0;
};

void hex_corner_rest(const double *corners, double *shape, double *hourglass);
void hex_material(float youngs, float poisson, double *shear, double *first);
double hex_cell_frequency(float youngs, float poisson, float density, double cell);
bool hex_make(HexElement &e, uint32_t cell, const uint32_t *nodes, const double *inv_mass,
              const double *shape, const double *hourglass, double volume, double shear,
              double first, double h, const HexDamping *damping = nullptr);
// lowestAngular (lowestmode.cc:31-203) of elements over nodes (mass, pinned
// flag, rest position): inverse iteration with a band L D L^T of the shifted
// stiffness; infinity without a free coordinate, 0 when a pivot vanishes
double hex_lowest_angular(const std::vector<HexElement> &elements, const std::vector<double> &mass,
                          const std::vector<uint8_t> &pinned, const std::vector<double> &rest);
// nodes by rest position along the longest extent, then the other axes,
// then id: the coordinate order of the lowest-mode band factor
std::vector<uint32_t> band_order(const std::vector<double> &rest, size_t nodes);
// rotationOf on the host: the quaternion iteration maximising tr(R^T F)
void hex_rotation_of(const double *F, const double *guess, uint32_t iterations, double *out);
// updateElement's rotation and capped strain rows at positions x (the
// element's quaternion advances, as updateElement's does)
void hex_strain(HexElement &e, const double *x, double *strain);
bool hex_fit_holds(const double *shape, const double *hourglass);

int build_hex_lattice(const talus_mi_lattice *lattice, const talus_mi_material *materials,
                      uint32_t lanes, const talus_mi_support *supports, uint32_t nsupports,
                      talus_mi_hex_mesh *out);
void hex_mesh_free(talus_mi_hex_mesh *mesh);

// One lattice body: mesh, materials, host state mirror and built elements.
struct LatticeBody {
  uint32_t nodes = 0, hexes = 0;
  std::vector<double> rest;       // 3*nodes
  std::vector<float> mass;        // nodes
  std::vector<uint8_t> anchored;  // nodes
  std::vector<uint32_t> hex;      // 8*hexes
  std::vector<float> volume;      // hexes
  std::vector<uint32_t> lane;     // hexes
  std::vector<talus_mi_material> materials;
  float voxel[3] = {};            // PreparedMesh::voxelSize (hex half extents are half of it)
  double cell = 0;                // smallest voxel side
  std::vector<uint8_t> node_colour, hex_colour;
  uint32_t substeps = 0;          // LatticeBody::substeps
  std::vector<double> x, v;       // host mirror of the state
  double h = 0;                   // substep the elements were built for (0 = none)
  std::vector<HexElement> elements;
  // plasticity (plasticflow.cc): per Gauss point the inverse plastic
  // deformation P (elastic = F P, row-major) and the plastic strain, empty
  // until the first flow; the P each cell's element was built from; and per
  // cell the rest its element was built on (volume 0 = none kept)
  bool yields = false;            // a lane has a finite yield stress
  std::vector<double> plastic;    // 72*hexes
  std::vector<double> hardening;  // 8*hexes
  std::vector<double> cached;     // 72*hexes
  std::vector<double> rest_shape, rest_hg, rest_vol;  // 9, 12, 1 per cell

  const talus_mi_material &material_of(uint32_t cell_index) const;
  bool pinned(size_t node) const;
  double stiffest() const;
  uint32_t island_substeps(double dt, uint32_t softstep) const;
  // builds the elements for substep h, seeded with the lattice turn of x
  // (restOf: the plastic rest, or the kept rest when its fit fails)
  void build_elements(double h);
  void ensure_plastic();  // identity P, zero strain
  // plasticRest, writeFitted (plasticflow.cc:268-340) of one cell
  void plastic_rest(uint32_t cell, double *shape, double *hourglass, double *volume) const;
  void write_fitted(uint32_t cell, const double *shape, const double *hourglass);
  // refreshFlowed (strainstep.cc:371-451): rebuilds the elements of the cells
  // whose P moved since their element was built; returns how many
  uint32_t refresh_flowed();

  // tearing (strainstep/tearing.cc): per-cell crack-face bits (bit
  // 2 axis + side; empty before the first tear) and whether the mesh changed
  // since the elements were built
  bool tears = false;             // a lane has a finite tear strain
  // damping (cellDamping, strainstep.cc:151-160): the lowest angular
  // frequency the elements were damped for (0 = not found yet)
  bool damped = false;            // a lane has a finite damping ratio
  double lowest = 0.0;
  HexDamping cell_damping(uint32_t cell, double h) const;
  // lowestAngular of the undamped elements on their rests
  double lowest_angular(double h) const;
  std::vector<uint8_t> crack;     // hexes
  bool retopo = false;
  // PreparedMesh::surfaceNodes as a per-node flag: buildMesh's (corners of
  // faces no other cell shares), found at the first tear or request, then
  // markCracks's additions; empty until then
  std::vector<uint8_t> surface;
  void ensure_surface();
  // 1 + the cell's lane tear strain (infinite when the lane never tears)
  double tear_limit(uint32_t cell) const;
  // stretchOf (tearing.cc:258-300) at the host positions: the largest
  // principal stretch over the eight corner frames and its direction (the
  // eigenvector's largest component positive)
  void cell_stretch(uint32_t cell, double *stretch, double *direction) const;
  // tearCells (tearing.cc:505-536) at the host positions; returns the node
  // copies made
  uint32_t tear_cells(const uint32_t *cells, uint32_t count);
  // latticePieces (resolution.cc:198-228)
  uint32_t pieces() const;
  // the element rebuild after a tear: every element is built again at the
  // same substep and keeps its cell's impulses and rotation
  // (strainstep.cc:610-623)
  void rebuild_torn();
};

// symmetric 3x3 eigen decomposition by cyclic Jacobi rotations (w: values,
// V: vectors as columns, row-major)
void sym_eigen3(const double *A, double *w, double *V);

// Fills a body from a hex mesh (makeLattice); returns a status.
int make_lattice(LatticeBody &body, const talus_mi_hex_mesh *mesh,
                 const talus_mi_material *materials, uint32_t lanes, const double *offset,
                 const double *velocity);

// Device arrays of every lattice body, stepped together (lattice.hip).
class LatticeSet {
 public:
  explicit LatticeSet(hipStream_t stream);
  ~LatticeSet();
  LatticeSet(const LatticeSet &) = delete;
  LatticeSet &operator=(const LatticeSet &) = delete;

  uint32_t add(LatticeBody &&body);
  uint32_t count() const { return uint32_t(bodies_.size()); }
  LatticeBody &body(uint32_t i);  // host mirror, current after download
  void set_state(uint32_t i, const double *x, const double *v);
  void get_state(uint32_t i, double *x, double *v);
  void set_substeps(uint32_t i, uint32_t substeps);
  void step(const talus_mi_lattice_params &params, talus_mi_lattice_report &report);
  // per Gauss point P and plastic strain (empty when the body never flowed)
  void get_plastic(uint32_t i, double *plastic, double *hardening, uint32_t *present);
  void set_plastic(uint32_t i, const double *plastic, const double *hardening);
  // tearCells / cellStretch at the body's current positions
  uint32_t tear(uint32_t i, const uint32_t *cells, uint32_t count);
  void cell_stretch(uint32_t i, uint32_t cell, double *stretch, double *direction);

 private:
  struct Device;
  void download();  // device state -> host mirrors
  // host mirrors -> device arrays, laid out for these per-body substep
  // counts and exact flags
  void upload(const std::vector<uint32_t> &substeps, const std::vector<uint8_t> &exact);
  hipStream_t stream_;
  std::vector<LatticeBody> bodies_;
  std::unique_ptr<Device> dev_;
  bool device_valid_ = false;  // device arrays hold the current state
  bool flowed_ = false;        // the last step flowed on the device: refresh before the next
  bool refresh_pending_ = false;  // P was set from the host
};

}  // namespace tmi
