/* field_oracle.h -- FP64 CPU restatement (test infrastructure only) of the
 * deformed-field re-bake of a hex lattice: the sampled SDF's probe
 * (src/talus/field/sampledsdf/sampledsdf.cc:234-320, 432-480), the deformed
 * probe (src/talus/deformable/contact/contact.cc:886-1425) and
 * bakeDeformedField (src/talus/deformable/deformedfield/deformedfield.cc).
 * Only tests/, smoke() and bench.py's CPU baseline use it, as the checker. */
#ifndef TALUS_FIELD_ORACLE_H
#define TALUS_FIELD_ORACLE_H

#include <stdint.h>

#include "hex_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SampledSdf (sampledsdf.h:19-32) without its brick summaries: samples
 * x-fastest, index x + res.x (y + res.y z). */
typedef struct orc_sdf {
  float origin[3];
  float voxel[3];
  uint32_t res[3];
  float *dist;
  uint8_t *mat;
} orc_sdf;

/* probe<double>(field, point).distance: trilinear inside the lattice, the
 * clamped point's value plus the distance to it outside. */
double orc_sdf_probe(const orc_sdf *field, const double *point);

/* field::distance (sampledsdf.cc:377-385) in float: the distance to the
 * lattice's box outside it, else the float trilinear sample. */
float orc_sdf_distance(const orc_sdf *field, const float *point);
/* surfaceThickness of collectSurface (solver.cc:342-351): per triangle
 * minus the mean float distance at its three rest positions. */
void orc_surface_thickness(const orc_sdf *field, const double *rest, uint32_t triangles, const uint32_t *tri,
                           float *thickness);

/* collectSurface's hex faces (solver.cc:252-303): each cell face split in
 * two triangles around its lowest node, kept when no other cell has it,
 * turned outward by the cell's opposite corner (float rest positions);
 * triangles in ascending order of their sorted triple with their cell and
 * lane, edges the sorted unique pairs.  Returns the triangle count; tri,
 * cell, lane take up to 12 per cell, edges up to 36 per cell. */
uint32_t orc_hex_surface(const orc_hex_mesh *mesh, uint32_t *tri, uint32_t *cell, uint32_t *lane,
                         uint32_t *edges, uint32_t *nedges);

/* probeDeformed (contact.cc:1287-1425) of one point against the lattice
 * (rest positions of the mesh, hex nodes, crack bits per hex or NULL) placed
 * at `position`: 1 and the distance, the element and its local coordinates
 * when found, else 0.  The probe is built with infinite reach (the bake's
 * DEFORMED_PAD cells of pad). */
int orc_probe_deformed(const orc_hex_mesh *mesh, const uint8_t *crack, const orc_sdf *rest,
                       const double *position, const double *point, double *distance, uint32_t *element,
                       double *local);

/* bakeDeformedField (deformedfield.cc:15-102): every sample of the rest
 * field's lattice, placed by the model [R | t] (3x4 row-major, FP64),
 * probes the deformed lattice; the found distance, else the rest sample (or
 * max(rest, pad) with carve), replaces current's sample when they differ by
 * more than 1e-7.  current's distances are rewritten in place; box_min and
 * box_ext (3 each) receive the changed samples' box (extent 0 when none).
 * Returns 0 when current is not on the rest field's lattice. */
int orc_bake_deformed(const orc_hex_mesh *mesh, const uint8_t *crack, const double *position, const double *model,
                      const orc_sdf *rest, orc_sdf *current, int carve, uint32_t *box_min, uint32_t *box_ext);

#ifdef __cplusplus
}
#endif

#endif
