/* contact_oracle.h -- FP32 CPU restatement (test infrastructure only) of the
 * deformable surface's contact generation against rigid partners' sampled
 * fields: findContacts (src/talus/deformable/contact/contact.cc:10-880) with
 * the float probe of a sampled field (src/talus/field/sampledsdf/
 * sampledsdf.cc:18-23, 61-93, 234-305, 387-402, 417-477).  Only tests/,
 * smoke() and bench.py's CPU baseline use it, as the checker. */
#ifndef TALUS_CONTACT_ORACLE_H
#define TALUS_CONTACT_ORACLE_H

#include <stdint.h>

#include "field_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ShapeProbe (contact.h:19-26); rotation w, x, y, z; lipschitz is
 * probeLipschitz of the field (+inf: no distance bound, nothing culled by it). */
typedef struct orc_shape {
  const orc_sdf *field;
  float center[3];
  float rotation[4];
  float offset[3];
  float linear[3];
  float angular[3];
  float lipschitz;
} orc_shape;

/* SurfaceContact (contact.h:64-74). */
typedef struct orc_surface_contact {
  uint32_t nodes[3];
  float weights[3];
  int32_t partner;
  float normal[3];
  float local_point[3];
  float gap;
  float bias;
  float friction;
  uint32_t triangle;
} orc_surface_contact;

/* The PreparedMesh arrays findContacts reads, plus the material lanes'
 * friction.  lane, thickness and the clusters may be NULL. */
typedef struct orc_contact_mesh {
  uint32_t nodes, triangles, lanes, clusters;
  const uint32_t *tri;
  const uint32_t *lane;
  const float *thickness;
  const uint32_t *cluster_faces;
  const uint32_t *cluster_starts; /* clusters + 1 */
  const uint8_t *anchored;
  float voxel[3];
  const float *friction; /* lanes */
} orc_contact_mesh;

/* What the culling did (no counterpart in the reference). */
typedef struct orc_contact_stats {
  uint32_t clusters_far;  /* (shape, cluster) pairs skipped by clusterFar */
  uint32_t faces_apart;   /* (shape, face) pairs whose swept boxes miss */
  uint32_t faces_far;     /* skipped by farFromShape */
  uint32_t faces_probed;  /* closestOnTriangle at the start */
  uint32_t candidates;    /* contacts before the duplicate test */
  uint32_t duplicates;
} orc_contact_stats;

/* findContacts: shapes in order, each shape's contacts in face order, a
 * contact dropped when its mesh point lies within 1e-2 of the largest voxel
 * side of an earlier contact of the same shape.  position and velocity are
 * the float state (3 per node).  Writes up to capacity contacts, returns
 * how many there are; stats may be NULL. */
uint32_t orc_find_contacts(const orc_contact_mesh *mesh, const float *position, const float *velocity,
                           const float *gravity, const orc_shape *shapes, uint32_t nshapes, float dt, float bias,
                           float threshold, int cull, orc_surface_contact *out, uint32_t capacity,
                           orc_contact_stats *stats);

/* probeLipschitz of a field whose bricks are all built (finishBricks):
 * sqrt(g^2 + 1), g the largest cellLipschitz over the lattice's cells. */
float orc_field_lipschitz(const orc_sdf *field);

/* probe<float>: distance and unit normal. */
void orc_field_probe(const orc_sdf *field, const float *point, float *distance, float *normal);
/* probeDistance and material (nearest sample's lane, 0 outside). */
float orc_field_probe_distance(const orc_sdf *field, const float *point);
uint32_t orc_field_material(const orc_sdf *field, const float *point);

/* barycentricOf (contact.cc:87-156): weights of the closest point of the
 * triangle (9 floats, corner-major) to point. */
void orc_barycentric(const float *triangle, const float *point, float *weights);

/* closestOnTriangle against a shape placed at its state after `elapsed`
 * (advanceShape): weights, point, distance and normal. */
void orc_closest_on_triangle(const orc_shape *shape, float elapsed, const float *triangle, float *weights,
                             float *point, float *distance, float *normal);

/* advanceShape (contact.cc:10-23): centre and rotation (w, x, y, z). */
void orc_advance_shape(const orc_shape *shape, float elapsed, float *center, float *rotation);

#ifdef __cplusplus
}
#endif

#endif
