/* screen_oracle.h -- FP64 CPU restatement (test infrastructure only) of the
 * FP64 stable neo-Hookean hex (src/talus/dynamics/neohookean/neohookean.cc),
 * the lattice energy ledger (src/talus/dynamics/ledger/ledger.cc:9-52) and
 * the screen factor of tier switching (src/talus/dynamics/screen/screen.cc).
 * Only tests/, smoke() and bench.py's CPU baseline use it, as the checker.
 *
 * Corners are 8 points, 3 doubles each, in the lattice corner order (corner
 * c at the cell's offsets (c & 1, c >> 1 & 1, c >> 2 & 1)); 3x3 matrices are
 * row-major. */
#ifndef TALUS_SCREEN_ORACLE_H
#define TALUS_SCREEN_ORACLE_H

#include <stdint.h>

#include "hex_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* moduliOf (neohookean.cc:82-90): shear mu and shifted lambda + mu, nu
 * clamped to [0, 0.49]. */
void orc_snh_moduli(float youngs, float poisson, double *shear, double *shifted);
/* cellEnergy (neohookean.cc:125-151): 8-point Gauss sum of
 * V/8 (mu/2 |H|^2 - mu (J - 1 - tr H) + shifted/2 (J - 1)^2). */
double orc_snh_cell_energy(const double *deformed, const double *rest, const double *half, double volume,
                           double shear, double shifted);
/* nodeTerms (neohookean.cc:153-196) of one corner: the energy gradient and
 * its own 3x3 Hessian block sum V/8 (mu |g|^2 I + shifted b b^T), b = cof(F) g. */
void orc_snh_node_terms(const double *deformed, const double *rest, const double *half, double volume,
                        double shear, double shifted, uint32_t corner, double *gradient, double *hessian);
/* restStiffness (neohookean.cc:198-229): 24x24 row-major, row 3c + r, column
 * 3o + k (block (c, o) = sum V/8 (mu (a.b) I + mu b a^T + (shifted - mu) a b^T),
 * a = g_c, b = g_o). */
void orc_snh_rest_stiffness(const double *half, double volume, double shear, double shifted, double *K);
/* volumeRatio (neohookean.cc:231-240): 1 + volumeChange at the cell centre. */
double orc_snh_volume_ratio(const double *deformed, const double *rest, const double *half);

/* latticeEnergy (ledger.cc:9-52) of a hex mesh at positions x, velocities v:
 * kinetic over free nodes, gravity -m g.x over every node, elastic the cell
 * energies with corner 0's displacement taken out of every corner. */
void orc_lattice_energy(const orc_hex_mesh *mesh, const double *shear, const double *shifted,
                        const double *x, const double *v, const double *gravity, double *kinetic,
                        double *potential, double *elastic);

/* RestBody (screen.h:17-25, restBodyOf screen.cc:26-83): node masses, total,
 * mass centre, inertia about it, principal axes (columns of a proper
 * rotation, eigenvalues ascending as moments) and the smallest rest extent. */
typedef struct orc_rest_body {
  double total;
  double centre[3];
  double inertia[9];
  double axes[9];
  double moments[3];
  double thickness;
} orc_rest_body;
void orc_rest_body_of(const orc_hex_mesh *mesh, orc_rest_body *out);

/* factorScreen + screenLoad (screen.cc:85-254) by a dense Cholesky factor:
 * the rest stiffness of every cell plus per-DOF shifts (1e6 x the largest
 * diagonal entry on held nodes: the mesh's anchored nodes and `held`; else
 * 1e-9 x largest / heaviest x node mass), the load relieved of its rigid
 * acceleration when nothing is held, the displacement with its mass-weighted
 * rigid part removed when nothing is held, dent = max |u| / thickness.
 * Returns 0 when the mesh is too large for the dense factor (3 nodes > 6000)
 * or the factor fails. */
int orc_screen_load(const orc_hex_mesh *mesh, const double *shear, const double *shifted,
                    const uint32_t *held, uint32_t nheld, const double *forces, double *displacement,
                    double *dent);

#ifdef __cplusplus
}
#endif

#endif
