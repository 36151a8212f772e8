/*
 * hex_oracle.h -- FP64 CPU restatement of the Talus strainstep hex-lattice
 * solve (the deformable path World::step runs today).
 *
 * TEST INFRASTRUCTURE ONLY, like talus_oracle.h: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Matrices are 3x3 row-major double[9].  Node vectors are xyz-interleaved.
 * A quaternion is double[4] = (w, x, y, z).
 */
#ifndef TALUS_HEX_ORACLE_H
#define TALUS_HEX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "talus_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_HEX_ROWS 18

/* One 18-row hex element (StrainElement, strainstep/element.h:40-73): six
 * corotated strain rows, twelve hourglass rows (four modes x three axes) and
 * one volume row solved through a Schur complement. */
typedef struct orc_hex {
  uint32_t cell;
  uint32_t nodes[8];
  double gradient[24];    /* per corner: shape^-T * signs / 8 */
  double inv_mass[8];
  double rest_hg[12];     /* rest hourglass vectors, four modes */
  double bend, size;
  double stiffness[ORC_HEX_ROWS], rate[ORC_HEX_ROWS], softness[ORC_HEX_ROWS],
      cap[ORC_HEX_ROWS];
  double inverse[ORC_HEX_ROWS * ORC_HEX_ROWS];
  double vol_stiffness, vol_rate, vol_softness, vol_cap;
  int volumetric;
  /* state */
  double impulse[ORC_HEX_ROWS], vol_impulse;
  double quat[4];
  double turn[9];
  double strain[ORC_HEX_ROWS];
  double vol_gradient[24];
  double coupling[ORC_HEX_ROWS], image[ORC_HEX_ROWS];
  double gain, rhs, schur;
} orc_hex;

double orc_hourglass_sign(uint32_t mode, uint32_t corner);
/* cornerRest (element.cc:199-214): rest shape and hourglass vectors of a
 * hex from its eight corners (corner c at offsets c&1, c>>1&1, c>>2&1). */
void orc_hex_corner_rest(const double *corners, double *shape, double *hourglass);
/* elementMaterial (element.cc:185-190): shear modulus and first Lame
 * parameter with nu clamped to [0, 0.499]. */
void orc_hex_material(float youngs, float poisson, double *shear, double *first);
/* cellFrequency (element.cc:192-197). */
double orc_hex_cell_frequency(float youngs, float poisson, float density, double cell);
/* makeElement (element.cc:244-342) for an undamped material (damping ratio
 * 1, theta 1); returns 0 when the reference returns no element. */
int orc_hex_make(orc_hex *e, uint32_t cell, const uint32_t *nodes, const double *inv_mass,
                 const double *shape, const double *hourglass, double volume, double shear,
                 double first, double h);
/* rotationOf (element.cc:344-372); F row-major, quaternions (w, x, y, z). */
void orc_rotation_of(const double *F, const double *guess, uint32_t iterations, double *out);
void orc_quat_matrix(const double *q, double *R);
void orc_hex_update(orc_hex *e, const double *positions);
void orc_hex_apply(const orc_hex *e, double *velocities, const double *impulse,
                   double vol_impulse);
void orc_hex_solve(orc_hex *e, double *velocities);

/* Hex mesh of a structured lattice (buildMesh with VolumeElementType::Hex,
 * solver.cc:2185-2316): first-touch nodes in a z, y, x scan, hex nodes in
 * corner order, volume = voxel product, node mass += rho V / 8 in float. */
typedef struct orc_hex_mesh {
  uint32_t nodes, hexes;
  double *rest;       /* 3*nodes, from the float samples */
  float *mass;        /* nodes */
  uint8_t *anchored;  /* nodes */
  uint32_t *hex;      /* 8*hexes */
  float *volume;      /* hexes */
  uint32_t *lane;     /* hexes */
  float voxel[3];
} orc_hex_mesh;

int orc_build_hex_lattice(const orc_lattice *lattice, const orc_material *materials,
                          uint32_t lanes, const orc_support *supports, uint32_t nsupports,
                          orc_hex_mesh *out);
void orc_hex_mesh_free(orc_hex_mesh *mesh);

/* A lattice body (LatticeBody, bodyset.h:22-46 and makeLattice,
 * bodyset.cc:13-86) with its cached elements (StrainLattice). */
typedef struct orc_lattice_body orc_lattice_body;

orc_lattice_body *orc_lattice_create(const orc_hex_mesh *mesh, const orc_material *materials,
                                     uint32_t lanes, const double *offset,
                                     const double *velocity);
void orc_lattice_free(orc_lattice_body *body);
void orc_lattice_set_substeps(orc_lattice_body *body, uint32_t substeps);
void orc_lattice_get_state(const orc_lattice_body *body, double *x, double *v);
void orc_lattice_set_state(orc_lattice_body *body, const double *x, const double *v);
/* Node colours of makeLattice (bodyset.cc:30-37): parity of the node's
 * lattice index per axis; a hex takes the colour of its corner 0. */
void orc_lattice_colours(const orc_lattice_body *body, uint8_t *node_colour,
                         uint8_t *hex_colour);
/* Elements built at the last step (cells without an element are skipped)
 * and their accumulated impulses (18 rows then the volume row). */
uint32_t orc_lattice_elements(const orc_lattice_body *body);
void orc_lattice_impulses(const orc_lattice_body *body, uint32_t *cells, double *impulses);
/* The substeps one step takes (islandSubsteps, strainstep.cc:551-567). */
uint32_t orc_lattice_substeps(const orc_lattice_body *body, double dt,
                              uint32_t softstep_substeps);
/* ---- lattice plasticity (plasticflow.cc, strainstep.cc:371-451) ---- */
/* shapeGradients (neohookean.cc:92-112): trilinear shape gradients at the
 * reference point `at` of a hex with half extents `half`, 3 per corner. */
void orc_hex_shape_gradients(const double *at, const double *half, double *grad);
/* returnMap (plasticflow.cc:110-155): radial return of the Hencky strain of
 * the elastic F to `yield`; returns 0 when the point stays elastic. */
int orc_hex_return_map(const double *elastic, double shear, double yield, double *mapped,
                       double *trial, double *strain);
/* fittedJacobian (plasticflow.cc:83-92) at Gauss point `point`. */
void orc_hex_fitted_jacobian(const double *shape, const double *hourglass, uint32_t point,
                             double *J);
/* Per Gauss point (8 per hex, point p at GAUSS_OFFSET * signs(p)) the
 * inverse plastic deformation P (elastic = F P, row-major 3x3) and the
 * accumulated plastic strain; empty until the first flow of a yielding body. */
uint32_t orc_lattice_has_plastic(const orc_lattice_body *body);
void orc_lattice_get_plastic(const orc_lattice_body *body, double *plastic, double *hardening);
void orc_lattice_set_plastic(orc_lattice_body *body, const double *plastic, const double *hardening);
/* flowLattice (plasticflow.cc:161-260) at the body's positions; returns the
 * largest change (one-norm of P) of any point. */
double orc_lattice_flow(orc_lattice_body *body);
/* plasticRest (plasticflow.cc:268-318), fitHolds (320-328), writeFitted
 * (330-340) of one cell. */
void orc_lattice_plastic_rest(const orc_lattice_body *body, uint32_t cell, double *shape,
                              double *hourglass, double *volume);
int orc_hex_fit_holds(const double *shape, const double *hourglass);
void orc_lattice_write_fitted(orc_lattice_body *body, uint32_t cell, const double *shape,
                              const double *hourglass);

/* Cell tearing (strainstep/tearing.cc).  cellStretch (tearing.cc:279-281
 * over stretchOf 258-300): the largest principal stretch of the cell's eight
 * corner frames and its direction (the eigenvector's sign: its largest
 * component positive).  overstretched (tearing.cc:488-503): the cells whose
 * stretch passes 1 + their lane's tear strain, written to cells, returns the
 * count.  tearCells (tearing.cc:505-536): cracks through each listed cell,
 * splits the nodes pinched by the cracks, recomputes the node masses and
 * marks the crack faces; returns the node copies made.  latticePieces
 * (resolution.cc:198-228): face-or-node-connected pieces of cells. */
void orc_lattice_cell_stretch(const orc_lattice_body *body, uint32_t cell, double *stretch,
                              double *direction);
uint32_t orc_lattice_overstretched(const orc_lattice_body *body, uint32_t *cells);
uint32_t orc_lattice_tear(orc_lattice_body *body, const uint32_t *cells, uint32_t count);
uint32_t orc_lattice_pieces(const orc_lattice_body *body);
/* PreparedMesh::surfaceNodes, ascending: buildMesh's (the corners of faces
 * no other cell shares), then markCracks's additions; returns the count and
 * writes the nodes when not NULL */
uint32_t orc_lattice_surface(orc_lattice_body *body, uint32_t *nodes);
/* the current node count, and (each when not null) the hex corner nodes
 * (8 per cell), node masses and per-cell crack-face bits (bit 2 axis + side,
 * zero before the first tear) */
uint32_t orc_lattice_topology(const orc_lattice_body *body, uint32_t *hex_nodes, float *node_mass,
                              uint8_t *crack_faces);
/* lowestAngular (lowestmode.cc:31-203) of the body's undamped elements on
 * their current rests (inverse iteration, band L D L^T of the shifted
 * stiffness): INFINITY when every node is pinned, 0 when the factor fails;
 * and the lowest frequency the body's elements were last damped for (0
 * before a damped body's first build).  Lanes with a finite damping ratio
 * damp each row toward it at that frequency (cellDamping, rowDamping). */
double orc_lattice_lowest_angular(const orc_lattice_body *body);
double orc_lattice_lowest(const orc_lattice_body *body);
/* symmetric 3x3 eigen decomposition by cyclic Jacobi rotations: eigenvalues
 * w and eigenvectors as the columns of V (row-major) */
void orc_sym_eigen3(const double *A, double *w, double *V);

/* One solveStrainstep of this body alone (strainstep.cc:569-842, no
 * contacts, joints or manifolds): the element rebuild after a tear (impulses
 * and rotations carried cell by cell) or refreshFlowed of the cells the last
 * flow moved, the substeps, then flowLattice when a lane yields and
 * overstretched / tearCells when a lane tears.
 * colored = 0 sweeps the elements in cell order as the reference does;
 * colored = 1 sweeps them colour by colour (hex colour 0..7), cells in order
 * within a colour.  Returns the substeps taken. */
uint32_t orc_lattice_step(orc_lattice_body *body, double dt, const double *gravity,
                          uint32_t softstep_substeps, uint32_t colored);

#ifdef __cplusplus
}
#endif

#endif
