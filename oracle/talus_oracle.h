/*
 * talus_oracle.h -- FP64 CPU restatement of the Talus deformable tet step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product (libtalus_mi.so) never links or calls it.
 *
 * Matrices are 3x3 row-major double[9].  Vectors of nodes are xyz-interleaved
 * double[3*n].  The rest-shape inverse Dm^-1 is stored row-major, so its rows
 * are the shape gradients g1, g2, g3 and g0 = -(g1+g2+g3)
 * (reference: src/talus/deformable/material/material.cc:199-207).
 */
#ifndef TALUS_ORACLE_H
#define TALUS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as talus::PhysicalMaterial
 * (src/talus/material.h:12-29); elastic_model 0 = Corotated, 1 = NeoHookean. */
typedef struct orc_material {
  float friction, restitution, density, strength, youngs_modulus,
      poissons_ratio, stiffness_damping, yield_stress, hardening_modulus;
  uint32_t elastic_model;
  float hysteresis, damping_ratio, tear_strain, plastic_max, crush_stress,
      crush_limit;
} orc_material;

/* Tet mesh (PreparedMesh, src/talus/mesh.h:25-44, tets only), SoA. */
typedef struct orc_mesh {
  uint32_t nodes, tets;
  double *rest;       /* 3*nodes, rest positions (from float lattice samples) */
  double *mass;       /* nodes, lumped rho*V/4 per incident tet */
  uint8_t *anchored;  /* nodes */
  uint32_t *tet;      /* 4*tets */
  double *dminv;      /* 9*tets, FP64 inverse of the rest edge matrix */
  double *volume;     /* tets */
  uint32_t *lane;     /* tets, material lane */
} orc_mesh;

/* Axis-aligned-or-not support volume (solver.h SupportVolume): a node is
 * anchored when |world_to_support * rest| <= half on every axis. */
typedef struct orc_support {
  float world_to_support[12]; /* row-major 3x4 affine */
  float half[3];
} orc_support;

/* Structured lattice input for the Kuhn mesh builder (buildMesh,
 * solver.cc:2185-2418, for fully solid cells). */
typedef struct orc_lattice {
  uint32_t res[3];          /* samples per axis (cells + 1) */
  float voxel[3];
  float origin[3];
  const uint8_t *solid;     /* cells, NULL = all solid */
  const uint32_t *cell_lane;/* cells, NULL = lane 0 */
  const float *sample_pos;  /* 3*samples or NULL = origin + voxel*index */
} orc_lattice;

typedef struct orc_settings {
  float dt;
  float gravity[3];
  uint32_t substeps;
  uint32_t cg_iterations;
  float tolerance;
  float spectral_radius;
  uint32_t threads;         /* 0/1 = serial */
  uint32_t converge;        /* corrections until the substep residual's root */
  uint32_t plastic_passes;  /* Settings::plasticPasses (solver.h:56); 0 = 40 */
} orc_settings;

typedef struct orc_report {
  uint32_t substeps;
  uint32_t cg_iterations;
  uint32_t cg_capped;
  uint32_t solves;
  double cg_worst_relative;
  double elastic_energy;
  double kinetic_energy;
  uint32_t plastic_growth;  /* any yield this step */
  uint32_t yield_events;    /* tets whose plastic state changed */
  uint32_t non_finite;
  /* the plastic Newton of bodies with yielding corotated lanes
   * (StepReport::plastic*, solver.cc:3951-4250) */
  uint32_t plastic_evaluations, plastic_solves, plastic_backtracks,
      plastic_linear_reads, plastic_handoffs, plastic_fallbacks,
      plastic_rejections, plastic_capped;
} orc_report;

typedef struct orc_body orc_body;

/* ---- mesh ---- */
int orc_build_lattice(const orc_lattice *lattice, const orc_material *materials,
                      uint32_t lanes, const orc_support *supports,
                      uint32_t nsupports, orc_mesh *out);
void orc_mesh_free(orc_mesh *mesh);

/* ---- materials ---- */
void orc_lame(float youngs, float poisson, double *lambda, double *mu);

/* ---- element kernels (per tet) ---- */
double orc_det3(const double *a);
void orc_polar(const double *F, double *R);
void orc_element_rotation(const double *F, double *R);
void orc_snh_tet(const double *u, const double *dminv, double volume,
                 double lambda, double mu, const double *plastic_offset,
                 double *force, double *energy);
void orc_corotated_tet(const double *x, const double *X, const double *dminv,
                       double volume, double lambda, double mu,
                       const double *plastic, double yield, double hardening,
                       double *R, double *force, double *energy,
                       double *plastic_next, uint32_t *yielding);
int orc_hencky_flow(const double *F, double mu, double yield, double hardening,
                    double plastic_max, double *plastic_offset,
                    double *alpha);
double orc_stretch(const double *F, double *direction);
void orc_eigen_sym3(const double *A, double *values, double *vectors);

/* ---- bodies ---- */
orc_body *orc_body_create(const orc_mesh *mesh, const orc_material *materials,
                          uint32_t lanes);
void orc_body_free(orc_body *body);
void orc_body_set_state(orc_body *body, const double *x, const double *v,
                        const double *a);
void orc_body_get_state(const orc_body *body, double *x, double *v, double *a);
void orc_body_get_rotations(const orc_body *body, double *R);
void orc_body_get_plastic(const orc_body *body, double *plastic);
void orc_body_set_plastic(orc_body *body, const double *plastic);
void orc_body_set_rotations(orc_body *body, const double *R);
/* Neo-Hookean volume frames cof(Fe) Fp^-T of the last evaluation (9 per tet,
 * identity for corotated tets); the operator's volumetric direction. */
void orc_body_get_frames(const orc_body *body, double *W);
void orc_body_set_frames(orc_body *body, const double *W);
int orc_step(orc_body *body, const orc_settings *settings, orc_report *report);

/* Per-tet forces (12 per tet) and energies at positions x, and the node
 * force sum, using the current plastic state. */
void orc_eval_forces(orc_body *body, const double *x, uint32_t threads,
                     double *tet_force, double *tet_energy, double *node_force,
                     double *rotations);
/* y = inertia*M*d + h^2 * sum_e R_e K_e R_e^T d_e, anchored rows zeroed
 * (applySystem, solver.cc:904-1048), with the body's committed rotations. */
void orc_apply_system(orc_body *body, double inertia, double step,
                      const double *d, double *y, uint32_t threads);
/* Per-tet stretch sqrt(lambda_max(F^T F)) at the body's positions. */
void orc_tet_stretch(orc_body *body, double *stretch);
void orc_stretch_tets(const uint32_t *tet, size_t tets, const double *rest, const double *x,
                      double *stretch);
/* J2 point law (returnMap, material.cc:594-630): stress and linearisation
 * lin11 = {theta, thetaBar, normal[9]}; tangentOf (636-647) applied to a
 * strain e. */
int orc_j2_point(const double *strain, const double *plastic, double lambda, double mu,
                 double yield, double hardening, double *stress, double *lin11);
void orc_j2_tangent(const double *lin11, double lambda, double mu, const double *e, double *out);
/* The consistent-tangent operator K d of a Newton attempt at positions x
 * (h = 1, no inertia); returns the loading tets. */
uint32_t orc_tangent_apply(orc_body *body, const double *x, const double *d, double *y);
/* The operator with the initial-stress (rotation-derivative) term at x:
 * y = sum over corotated tets of (R K R^T + initialStress) d, h = 1. */
void orc_stressed_apply(orc_body *body, const double *x, const double *d, double *y);
/* The operator with stiffness damping at x: inertia M d + step^2 K d +
 * damp sum_e tau_e P R K R^T P d (P the element's rigid-mode removal). */
void orc_damped_apply(orc_body *body, const double *x, double inertia, double step, double damp,
                      const double *d, double *y);
/* Yield flags of the last step (1 per tet). */
void orc_yield_flags(const orc_body *body, uint8_t *flags);

/* Tear topology (tearCells, tearing.cc:485-513, restated for tets) on plain
 * arrays with room for `capacity` nodes: tet (4*tets, in/out), rest and x
 * (3*capacity, in/out), anchored (capacity, in/out); torn tets ascending.
 * Appends split nodes (origin[i] = node that node nodes+i came from) and
 * returns the new node count. */
uint32_t orc_tear_topology(uint32_t nodes, uint32_t tets, uint32_t *tet, double *rest,
                           double *x, uint8_t *anchored, uint32_t capacity,
                           const uint32_t *torn, uint32_t ntorn, uint32_t *origin);
/* Tears a body at its current positions; returns the number of new nodes
 * (state of new nodes copied from their origin, masses recomputed). */
uint32_t orc_body_tear(orc_body *body, const uint32_t *torn, uint32_t ntorn);
uint32_t orc_body_nodes(const orc_body *body);
void orc_body_topology(const orc_body *body, uint32_t *tet, double *rest, double *mass,
                       uint8_t *anchored);

/* Per-element Lame parameters (FP64, overriding the material lanes) and
 * FP64 rest positions (rest inverse, volumes and lumped masses recomputed;
 * -1 if a tet degenerates or inverts). */
void orc_body_set_lame(orc_body *body, const double *lambda, const double *mu);
void orc_body_get_lame(const orc_body *body, double *lambda, double *mu);
int orc_body_set_rest(orc_body *body, const double *X);

/* Differentiable rollout (no counterpart in the reference): `steps`
 * converged steps (corrections to the root of every substep's residual),
 * checkpointing x, v, a every `every` steps; then the adjoint of a loss with
 * gradients x_bar, v_bar (3 per node) at the last state: per-tet dL/dlambda,
 * dL/dmu, per-node dL/dX (rest), and the initial-state gradients.  Elastic
 * neo-Hookean bodies only (-2 otherwise).  The body's state is left at the
 * end of the rollout. */
int orc_rollout(orc_body *body, const orc_settings *settings, uint32_t steps, uint32_t every);
int orc_rollout_backward(orc_body *body, const double *x_bar, const double *v_bar,
                         double *lam_bar, double *mu_bar, double *X_bar, double *x0_bar,
                         double *v0_bar, double *a0_bar, double tolerance);
/* y = K(pos) d with K the neo-Hookean Hessian of the elastic energy. */
void orc_hessian_apply(orc_body *body, const double *pos, const double *d, double *y);

/* clusterSurface (contact.cc:540-597): the ntri triangles (3 nodes each)
 * in Morton order of their rest centres (float arithmetic on the float rest
 * samples), starts every 32 and the count last; returns the clusters. */
uint32_t orc_cluster_surface(const uint32_t *tri, uint32_t ntri, const double *rest, uint32_t *faces,
                             uint32_t *starts);
/* Boundary surface and crack faces of a tet mesh (see talus_oracle.c). */
uint32_t orc_surface(uint32_t nodes, uint32_t tets, const uint32_t *tet, const uint32_t *lane,
                     const double *rest, uint32_t *tri, uint32_t *tri_tet, uint32_t *tri_lane,
                     uint32_t *edges, uint32_t *edges_out, uint32_t *surf, uint32_t *nodes_out,
                     uint8_t *crack);

#ifdef __cplusplus
}
#endif

#endif
