/*
 * talus_mi.h -- C ABI of libtalus_mi.so, the MI355X (gfx950) deformable tet
 * step behind talus::World::step() and talus::PhysicalMaterial.
 *
 * Plain C: pointers and sizes only, no exceptions across the boundary.  Every
 * call returns TALUS_MI_OK (0) or a negative status; talus_mi_last_error()
 * gives the message of the calling thread's last failure.  One context owns
 * one HIP stream and all device buffers; calls on a context must be
 * serialized by the caller (src/talus README: "Access to a world must be
 * serialized").  Results are bit-reproducible run to run: assembly sums in a
 * fixed order and no floating-point atomics are used.
 *
 * Matrices are 3x3 row-major; node vectors are xyz-interleaved.
 */
#ifndef TALUS_MI_H
#define TALUS_MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TALUS_MI_OK 0
#define TALUS_MI_ERR_ARGUMENT -1
#define TALUS_MI_ERR_DEVICE -2
#define TALUS_MI_ERR_STATE -3
#define TALUS_MI_ERR_MEMORY -4

/* Arithmetic of the element kernels and the inner solve. */
#define TALUS_MI_FP64 0u  /* plain FP64 everywhere */
#define TALUS_MI_MIXED 1u /* FP32 elements/PCG, FP64 Dm^-1, pivots, residuals */

/* Preconditioner of the implicit solve. */
#define TALUS_MI_BLOCK_JACOBI 0u /* node 3x3 diagonal inverse (solver.cc:1574) */
#define TALUS_MI_FACTOR 1u       /* supernodal block Cholesky (factor.cc:549) */

/* Step flags. */
#define TALUS_MI_DETECT_TEARS 1u /* tear events and host topology update */
#define TALUS_MI_RECORD_EVENTS 2u
/* corrections repeated until the increment is below 1e-12 of the
 * acceleration in every body (at most 63): each substep lands on the root of
 * its residual, as the differentiable rollout requires */
#define TALUS_MI_CONVERGE 4u
/* launch every PCG kernel directly instead of replaying the captured HIP
 * graph of one iteration (for comparison; results are identical) */
#define TALUS_MI_DIRECT_LAUNCH 8u
/* apply the block Cholesky in the frame it was factored in: no bodyRotation
 * (factor.cc:615-646; by default the factor follows each body's best-fit
 * rotation since its factorization, as the reference's) */
#define TALUS_MI_FIXED_FRAME 32u

/* elastic_model (ElasticModel, src/talus/material.h:8-11) */
#define TALUS_MI_COROTATED 0u
#define TALUS_MI_NEO_HOOKEAN 1u

typedef struct talus_mi_ctx talus_mi_ctx;

/* Same fields, order and meaning as talus::PhysicalMaterial
 * (src/talus/material.h:12-29).  elastic_model: 0 Corotated, 1 NeoHookean.
 * Lanes beyond the array use a default-constructed material
 * (laneOf, src/talus/deformable/solver/solver.cc:30-37). */
typedef struct talus_mi_material {
  float friction;
  float restitution;
  float density;
  float strength;
  float youngs_modulus;
  float poissons_ratio;
  float stiffness_damping;
  float yield_stress;
  float hardening_modulus;
  uint32_t elastic_model;
  float hysteresis;
  float damping_ratio;
  float tear_strain;
  float plastic_max;
  float crush_stress;
  float crush_limit;
} talus_mi_material;

/* Host SoA view of a tet PreparedMesh (src/talus/mesh.h:11-44).  Arrays are
 * caller-owned and copied by talus_mi_body_add. */
typedef struct talus_mi_mesh {
  uint32_t nodes;
  uint32_t tets;
  float *rest_positions; /* 3*nodes */
  float *node_mass;      /* nodes; NULL on input = rho*V/4 per tet in FP64 */
  uint8_t *anchored;     /* nodes */
  uint32_t *tet_nodes;   /* 4*tets, positively oriented */
  float *rest_inverse;   /* 9*tets; output of the builder, not read on input */
  float *volume;         /* tets;   output of the builder, not read on input */
  uint32_t *tet_material;/* tets, material lane */
} talus_mi_mesh;

/* Structured lattice for the Kuhn tet builder (buildMesh, solver.cc:2185). */
typedef struct talus_mi_lattice {
  uint32_t resolution[3];   /* samples per axis = cells + 1 */
  float voxel_size[3];
  float origin[3];
  const uint8_t *solid;     /* per cell, NULL = all solid (fill 1) */
  const uint32_t *cell_lane;/* per cell, NULL = lane 0 */
  const float *sample_positions; /* 3 per sample, NULL = origin+voxel*index */
} talus_mi_lattice;

/* SupportVolume (solver.h): anchored iff |world_to_support * p| <= half. */
typedef struct talus_mi_support {
  float world_to_support[12]; /* row-major 3x4 affine */
  float half_extents[3];
} talus_mi_support;

/* Step parameters: deformable::Settings (solver.h:51-107) subset plus the
 * IntegrationParameters dt and gravity (integrationparameters.h:9-33). */
typedef struct talus_mi_step_params {
  float dt;
  float gravity[3];
  uint32_t substeps;        /* default 1 */
  uint32_t cg_iterations;   /* default 64 */
  uint32_t factor_budget;   /* default 32 (cap with a factor) */
  float tolerance;          /* default 1e-3, relative residual */
  float spectral_radius;    /* default 0.5 */
  uint32_t preconditioner;  /* TALUS_MI_BLOCK_JACOBI / TALUS_MI_FACTOR */
  uint32_t flags;           /* TALUS_MI_DETECT_TEARS | TALUS_MI_RECORD_EVENTS */
  uint32_t plastic_passes;  /* Settings::plasticPasses (solver.h:56), default
                               40: Newton attempts per substep of a body with
                               yielding corotated lanes; 1 = the elastic
                               three-solve scheme */
  /* FP64 iterative refinement of every elastic solve (no counterpart in the
   * reference, whose CPU solve is FP32): 0 = off (default).  Otherwise,
   * after each solve the residual right - A a is formed in FP64 (the
   * operator in FP64 arithmetic on the FP64 rest data and acceleration) and
   * correction solves A d = residual (the PCG at `tolerance` relative to the
   * residual) are added until |right - A a| <= refine_tolerance |right| in
   * every body or refine_rounds corrections were made.  Meant for the mixed
   * path at tolerances below what FP32 vectors reach (about 1e-6). */
  float refine_tolerance;
  uint32_t refine_rounds;   /* default 8 */
} talus_mi_step_params;

/* deformable::StepReport subset (solver.h:211-311); filled per call. */
typedef struct talus_mi_report {
  uint32_t bodies;
  uint32_t nodes;
  uint32_t elements;
  uint32_t substeps;        /* largest over bodies */
  uint32_t cg_iterations;   /* summed over bodies and solves */
  uint32_t cg_capped;
  uint32_t solves;
  uint32_t refactors;
  double cg_worst_relative;
  double elastic_energy;
  double kinetic_energy;
  uint32_t plastic_growth;
  uint32_t yield_events;
  uint32_t tear_events;
  uint32_t split_nodes;
  uint32_t non_finite;
  uint32_t preconditioner_applications; /* block Cholesky solves this step */
  double device_seconds;    /* hipEvent time of the whole step */
  /* the plastic Newton (StepReport::plastic*, solver.h:211-311), summed over
   * bodies and substeps: assemblies, solves, halved increments, elastic
   * reads taken instead, hand-backs to Newton, accepted elastic fallbacks,
   * rejected fallbacks (lowest-residual iterate kept), unconverged substeps */
  uint32_t plastic_evaluations;
  uint32_t plastic_solves;
  uint32_t plastic_backtracks;
  uint32_t plastic_linear_reads;
  uint32_t plastic_handoffs;
  uint32_t plastic_fallbacks;
  uint32_t plastic_rejections;
  uint32_t plastic_capped;
  /* FP64 iterative refinement (refine_tolerance > 0): correction solves,
   * summed over bodies and solves, and the largest final FP64 relative
   * residual |right - A a| / |right| over bodies and solves */
  uint32_t refinements;
  uint32_t refine_unmet;    /* solves of a body left above refine_tolerance */
  double refine_worst_relative;
} talus_mi_report;

/* The last talus_mi_step's StepReport of one body alone (deformable::step
 * reports per body, solver.h:211-311): nodes, elements, substeps, CG
 * iterations, capped and total solves, worst relative residual, energies,
 * yield events, non-finite flag and the plastic Newton's counters;
 * refactors, preconditioner applications and device time are the batch's
 * and stay zero here.  All zero before the first step and after an edit. */
int talus_mi_body_report(talus_mi_ctx *ctx, uint32_t body, talus_mi_report *out);

#define TALUS_MI_EVENT_YIELD 0u
#define TALUS_MI_EVENT_TEAR 1u

typedef struct talus_mi_event {
  uint32_t kind;
  uint32_t body;
  uint32_t tet;
  uint32_t step;
} talus_mi_event;

/* ---- context ---- */
/* Creates a context on HIP device `device` computing in `precision`
 * (TALUS_MI_FP64 / TALUS_MI_MIXED). */
int talus_mi_create(talus_mi_ctx **out, int device, uint32_t precision);
int talus_mi_destroy(talus_mi_ctx *ctx);
const char *talus_mi_last_error(void);
void talus_mi_default_params(talus_mi_step_params *params);

/* ---- mesh (replaces deformable::buildMesh, solver.cc:2185-2418) ---- */
/* Kuhn split of every solid cell into 6 tets, nodes numbered in first-touch
 * z-y-x cell order; arrays are malloc'ed, release with talus_mi_mesh_free. */
int talus_mi_build_lattice(const talus_mi_lattice *lattice,
                           const talus_mi_material *materials, uint32_t lanes,
                           const talus_mi_support *supports,
                           uint32_t nsupports, talus_mi_mesh *out);
void talus_mi_mesh_free(talus_mi_mesh *mesh);

/* Boundary surface of a tet mesh and its crack faces (the tet faces of
 * collectSurface, src/talus/deformable/solver/solver.cc:220-364, and
 * markCracks, src/talus/dynamics/strainstep/tearing.cc:406-461, for tets).
 * A face (corner triple {0,1,2}, {0,1,3}, {0,2,3} or {1,2,3} of a tet) is on
 * the surface when exactly one tet has it; its triangle keeps the tet's
 * corner order, flipped to face away from the tet.  Triangles ascend by
 * their sorted node triple; edges are unique (low, high) pairs in ascending
 * order; surface nodes ascend.  crack_faces[e] has bit f set when face f of
 * tet e is a crack face: a surface face whose three rest positions are
 * those of another surface face (the two sides of a tear).  Arrays are
 * malloc'ed; release with talus_mi_surface_free. */
typedef struct talus_mi_surface {
  uint32_t triangles, edges, nodes, tets;
  uint32_t *triangle_nodes; /* 3*triangles, outward */
  uint32_t *triangle_tet;   /* triangles */
  uint32_t *triangle_lane;  /* triangles: the tet's material lane */
  uint32_t *edge_nodes;     /* 2*edges */
  uint32_t *surface_nodes;  /* nodes */
  uint8_t *crack_faces;     /* tets */
  /* clusterSurface (src/talus/deformable/contact/contact.cc:540-597):
   * triangle indices in Morton order of their rest centres (10 bits per
   * axis over the centres' box), cut into clusters of 32; cluster c holds
   * cluster_faces[cluster_starts[c] .. cluster_starts[c + 1]) */
  uint32_t clusters;
  uint32_t *cluster_faces;  /* triangles */
  uint32_t *cluster_starts; /* clusters + 1 */
} talus_mi_surface;
int talus_mi_mesh_surface(const talus_mi_mesh *mesh, talus_mi_surface *out);
void talus_mi_surface_free(talus_mi_surface *surface);

struct talus_mi_sdf; /* the sampled field, defined with the lattices below */
/* surfaceThickness of collectSurface (src/talus/deformable/solver/
 * solver.cc:342-351): per triangle (3 node indices each) minus the mean of
 * the lattice field's distance (field::distance, sampledsdf.cc:377-385, in
 * float) at its three rest positions (3 per node). */
int talus_mi_surface_thickness(const struct talus_mi_sdf *lattice, const double *rest, uint32_t triangles,
                               const uint32_t *triangle_nodes, float *thickness);

/* Tear topology update (tearCells, src/talus/dynamics/strainstep/
 * tearing.cc:485-513, on tets) of a host mesh at positions x (3 per node):
 * nodes around each torn tet (ascending indices) are split, new nodes are
 * appended, masses recomputed as rho*V/4.  *out is malloc'ed (release with
 * talus_mi_mesh_free); *origin (one entry per new node: the node it was
 * split from) is malloc'ed, release with talus_mi_free. */
int talus_mi_tear_mesh(const talus_mi_mesh *mesh, const talus_mi_material *materials,
                       uint32_t lanes, const double *x, const uint32_t *torn,
                       uint32_t ntorn, talus_mi_mesh *out, uint32_t **origin,
                       uint32_t *new_nodes);
void talus_mi_free(void *p);

/* ---- bodies (the tet body set stepped inside World::step) ---- */
int talus_mi_body_add(talus_mi_ctx *ctx, const talus_mi_mesh *mesh,
                      const talus_mi_material *materials, uint32_t lanes,
                      uint32_t *body);
int talus_mi_body_info(talus_mi_ctx *ctx, uint32_t body, uint32_t *nodes,
                       uint32_t *tets);
/* DeformableState position/velocity/acceleration (mesh.h:56-62); NULL skips */
int talus_mi_state_set(talus_mi_ctx *ctx, uint32_t body, const double *x,
                       const double *v, const double *a);
int talus_mi_state_get(talus_mi_ctx *ctx, uint32_t body, double *x, double *v,
                       double *a);
/* committed[e].rotation (9 per tet) and plastic points (10 per tet:
 * corotated eps_p[9] + alpha, neo-Hookean (Fp^-1 - I)[9] + alpha) */
int talus_mi_rotations_get(talus_mi_ctx *ctx, uint32_t body, double *R);
int talus_mi_rotations_set(talus_mi_ctx *ctx, uint32_t body, const double *R);
/* Neo-Hookean volume frames W = cof(Fe) Fp^-T (9 per tet, identity for
 * corotated tets) of the last element evaluation: the operator's volumetric
 * direction for neo-Hookean tets (no counterpart in the reference, whose tet
 * path is corotated; kept with the rotations so a restored state steps
 * identically). */
int talus_mi_frames_get(talus_mi_ctx *ctx, uint32_t body, double *W);
int talus_mi_frames_set(talus_mi_ctx *ctx, uint32_t body, const double *W);
int talus_mi_plastic_get(talus_mi_ctx *ctx, uint32_t body, double *plastic);
int talus_mi_plastic_set(talus_mi_ctx *ctx, uint32_t body,
                         const double *plastic);
/* Current connectivity, rest positions and masses (they change on tears). */
int talus_mi_topology_get(talus_mi_ctx *ctx, uint32_t body,
                          uint32_t *tet_nodes, double *rest, double *mass,
                          uint8_t *anchored);

/* ---- stepping (deformable::step, solver.cc:3472; World::step) ---- */
int talus_mi_step(talus_mi_ctx *ctx, const talus_mi_step_params *params,
                  talus_mi_report *report);
/* Yield and tear events recorded since the last call (ascending tet order per
 * body and step). */
int talus_mi_events(talus_mi_ctx *ctx, talus_mi_event *out, uint32_t capacity,
                    uint32_t *count);

/* ---- kernel-level entry points (parity and roofline) ---- */
/* Element kernel alone at positions x: per-tet corner forces (12 per tet),
 * energies, assembled node forces, committed rotations (each may be NULL). */
int talus_mi_eval_forces(talus_mi_ctx *ctx, uint32_t body, const double *x,
                         double *tet_forces, double *tet_energy,
                         double *node_forces, double *rotations);
/* Matrix-free operator y = inertia*M*d + step^2 sum_e R K R^T d (applySystem,
 * solver.cc:904-1048) with the committed rotations. */
int talus_mi_apply_system(talus_mi_ctx *ctx, uint32_t body, double inertia,
                          double step, const double *d, double *y);
/* The operator with the initial-stress (rotation-derivative) term of the
 * static Newton (scratch.stressed, solver.cc:2623; initialStress and
 * applyInitialStress, material.cc:1077-1174): at positions x of the body
 * (which become its evaluation state: rotations and forces), y = inertia*M*d
 * + step^2 sum_e (R K R^T + initialStress_e) d over its corotated tets
 * (neo-Hookean tets add their operator only), anchored rows zero. */
int talus_mi_apply_system_stressed(talus_mi_ctx *ctx, uint32_t body, const double *x, double inertia,
                                   double step, const double *d, double *y);
/* The operator with the stiffness-damping term (applySystem's addDamping,
 * solver.cc:1006-1007; applyDamping, factor/factor.cc:848-857): at
 * positions x of the body (which become its evaluation state: rotations and
 * rigid fits), y = inertia*M*d + step^2 sum_e R K R^T d + damp sum_e tau_e
 * P_e R K R^T P_e d, tau_e the tet's stiffness_damping and P_e the removal
 * of its corners' rigid motion; anchored rows zero.  The step's damping
 * coefficient is (1 - alpha_f) gamma h (solver.cc:3590). */
int talus_mi_apply_system_damped(talus_mi_ctx *ctx, uint32_t body, const double *x, double inertia,
                                 double step, double damp, const double *d, double *y);
/* Size of the allocation-free path's per-GPU scratch array (MiB). */
#define TALUS_MI_NOALLOC_SCRATCH_MIB 48
/* talus_mi_apply_system without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  d and y are DEVICE pointers to 3*nodes doubles of the body;
 * the operator is enqueued on talus_mi_stream(ctx) and y holds the same
 * values as talus_mi_apply_system's once that stream has synchronized.  The
 * body's per-tet slots (12*tets values in the context's precision) live in
 * one static device array of TALUS_MI_NOALLOC_SCRATCH_MIB MiB per GPU,
 * resolved by talus_mi_create: bodies beyond it (more than 524,288 tets in
 * FP64, 1,048,576 in mixed precision) return TALUS_MI_ERR_ARGUMENT.  The
 * context must be built since its last body edit (talus_mi_step,
 * talus_mi_apply_system, ...), else TALUS_MI_ERR_STATE.  Every call on a GPU
 * shares that GPU's array and keeps no state in it between calls: calls that
 * overlap on one GPU (other streams or contexts) race. */
int talus_mi_apply_system_noalloc(talus_mi_ctx *ctx, uint32_t body, double inertia,
                                  double step, const double *d, double *y);
/* Times one kernel at the current state with HIP events on the context
 * stream (`repeats` rounds after one warm-up) and states its algorithmic
 * HBM bytes (what the algorithm must move: element arrays once per tet, node
 * arrays once per node, factor entries once).  Kinds: 0 element assembly,
 * 1 operator element pass, 2 operator node gather (one launch over every
 * body each); 3 / 4 forward / backward solve, one-lane-per-row shape,
 * 5 / 6 the rows-per-wavefront shape, 7 / 8 the split shape (one launch per
 * elimination level that uses the shape under the current plan).  avg_ms and bytes are per launch;
 * launches = launches of the kernel per timed unit (per preconditioner
 * application for kinds 3-6, else 1). */
typedef struct talus_mi_kernel_timing {
  double avg_ms;
  double bytes;
  uint32_t launches;
  uint32_t reserved;
} talus_mi_kernel_timing;
int talus_mi_time_kernel(talus_mi_ctx *ctx, uint32_t kind, uint32_t repeats,
                         talus_mi_kernel_timing *out);
/* Block Cholesky preconditioner alone (SystemFactor, factor.cc:549-704):
 * factorizes inertia*M + step^2 sum_e R K R^T with the committed rotations
 * (anchored rows removed) and returns z = (L L^T)^-1 r for one body. */
int talus_mi_factor_solve(talus_mi_ctx *ctx, uint32_t body, double inertia, double step,
                          const double *r, double *z);

typedef struct talus_mi_factor_info {
  uint32_t fronts;
  uint32_t levels;
  uint32_t ok;             /* 0 after a non-positive pivot (block Jacobi used) */
  uint32_t reserved;
  uint64_t entries;        /* stored factor entries (DOF) */
  double flops;            /* dense supernodal factorization flops */
  double factor_seconds;   /* device time of the last factorization */
} talus_mi_factor_info;

/* bodyRotation (factor.cc:615-646) of a body at its current positions: the
 * polar factor of sum_n m_n (x_n - c)(r_n - c_r)^T with r the positions at
 * its last factorization, row-major; the identity before a factorization. */
int talus_mi_body_frame_get(talus_mi_ctx *ctx, uint32_t body, double *frame);

/* Statistics of the current factor (zeros before the first factorization). */
int talus_mi_factor_info_get(talus_mi_ctx *ctx, talus_mi_factor_info *info);

/* Kernels of one factorization (SystemFactor::refresh, factor.cc:549-613,
 * here the supernodal block Cholesky): element-pair assembly, extend-add,
 * 64x64 diagonal Cholesky, panel TRSM, trailing SYRK/GEMM, L11^-1 and the
 * packing of G = [L11^-1; L21 L11^-1]. */
#define TALUS_MI_FK_ASSEMBLE 0u
#define TALUS_MI_FK_EXTEND_ADD 1u
#define TALUS_MI_FK_POTRF 2u
#define TALUS_MI_FK_TRSM 3u
#define TALUS_MI_FK_SYRK 4u
#define TALUS_MI_FK_TRTRI 5u
#define TALUS_MI_FK_PACK 6u
typedef struct talus_mi_factor_kernel {
  uint32_t kind;        /* TALUS_MI_FK_* */
  uint32_t launches;
  uint64_t workgroups;
  double mfma_flops;    /* FP64 flops of the 64x64x64 MFMA tiles issued */
  double useful_flops;  /* flops on the fronts' own entries (tiles clipped) */
  double ms;            /* summed HIP-event time of the launches */
} talus_mi_factor_kernel;
/* Factors every body once (with the current step keys, or h = 1 and unit
 * inertia before the first step) with a HIP event pair around every launch
 * and returns one record per kernel kind, summed over its launches. */
int talus_mi_factor_profile(talus_mi_ctx *ctx, talus_mi_factor_kernel *out, uint32_t capacity,
                            uint32_t *count);

/* ---- differentiation (no counterpart in the reference: the adjoint of
 * the implicit rollout, for gradient-based identification and design) ---- */

/* Per-tet Lame parameters in FP64, overriding the lanes' E and nu for this
 * body until the next talus_mi_lame_set (tears keep them).  Gradients of
 * E and nu follow from these by the chain rule of the Lame map. */
int talus_mi_lame_set(talus_mi_ctx *ctx, uint32_t body, const double *lambda, const double *mu);
int talus_mi_lame_get(talus_mi_ctx *ctx, uint32_t body, double *lambda, double *mu);
/* FP64 rest positions (3 per node): rest inverse, volumes and (for density
 * masses) the lumped masses are recomputed; the state is kept. */
int talus_mi_rest_set(talus_mi_ctx *ctx, uint32_t body, const double *rest);
/* `steps` converged steps (TALUS_MI_CONVERGE, tears off) from the current
 * state, keeping x, v, a of every body every `every` steps on the host. */
int talus_mi_rollout(talus_mi_ctx *ctx, const talus_mi_step_params *params, uint32_t steps,
                     uint32_t every, talus_mi_report *report);
/* Reverse pass of the last rollout for one body, for a loss with gradients
 * x_bar, v_bar (3 per node; v_bar may be NULL) at its final state: per-tet
 * dL/dlambda, dL/dmu, per-node dL/drest and dL/dx0, dL/dv0, dL/da0 of the
 * rollout's initial state (NULL outputs skipped).  Segments are recomputed
 * from the checkpoints; every substep's adjoint system (inertia M + h_eff^2
 * K(x_w)) is solved by PCG to `tolerance` with the step's preconditioner.
 * Elastic neo-Hookean bodies only; the state is left at the rollout's end. */
int talus_mi_rollout_backward(talus_mi_ctx *ctx, uint32_t body, const double *x_bar,
                              const double *v_bar, double tolerance, double *lambda_bar,
                              double *mu_bar, double *rest_bar, double *x0_bar, double *v0_bar,
                              double *a0_bar);

/* Autotuner of the preconditioner solve (no counterpart in the reference,
 * whose factor solve is Eigen's).  The solve plan holds, per elimination
 * level, the kernel shape of the forward and of the backward pass (2 bytes
 * per level: 0 = rows per wavefront; for the fronts with few pivots 1 = one
 * row/column per lane, 2 = 64 rows/columns per workgroup with the other
 * dimension split over its wavefronts).  The factor's analysis sets a default plan from
 * the front shapes; talus_mi_tune times both shapes of every level on the
 * device and keeps the faster.  The plan changes the summation order of the
 * solve: results are bit-identical for a fixed plan, so a caller that needs
 * run-to-run identity tunes once and re-applies the plan with
 * talus_mi_plan_set.  A new analysis (bodies added, tears) restores the
 * default plan. */
typedef struct talus_mi_tune_report {
  uint32_t levels;
  uint32_t lanes_forward;  /* levels whose forward pass uses the lane shape */
  uint32_t lanes_backward;
  uint32_t leaf;           /* supernode leaf size in use */
  double before_seconds;   /* one preconditioner application, previous plan */
  double after_seconds;    /* ... tuned plan */
} talus_mi_tune_report;

int talus_mi_tune(talus_mi_ctx *ctx, uint32_t repeats, talus_mi_tune_report *report);
/* Supernode leaf size of the block Cholesky preconditioner: the nested
 * dissection stops at fronts of at most `leaf` nodes (default 64).  Smaller
 * leaves give less fill (fewer bytes per preconditioner application) and
 * more elimination levels.  Changing it rebuilds the context's device
 * data at the next call (the state is kept) and resets the solve plan.  No
 * counterpart in the reference, whose SimplicialLLT has no supernodes. */
#define TALUS_MI_LEAF_MIN 8u
#define TALUS_MI_LEAF_MAX 512u
int talus_mi_set_leaf(talus_mi_ctx *ctx, uint32_t leaf);
int talus_mi_leaf_get(talus_mi_ctx *ctx, uint32_t *leaf);
/* The autotuner over the leaf size: for each candidate, rebuild, factorize
 * and tune the per-level solve kernels (talus_mi_tune), then keep the leaf
 * whose tuned preconditioner application is fastest (HIP events, best of
 * `repeats`).  times[i] (when not null) receives candidate i's application
 * time in seconds; the report is the chosen leaf's talus_mi_tune report,
 * with before_seconds the application time before the call. */
int talus_mi_tune_leaf(talus_mi_ctx *ctx, const uint32_t *candidates, uint32_t count, uint32_t repeats,
                       double *times, talus_mi_tune_report *report);
/* Copies the plan (2*levels bytes) when capacity allows; *count = its size. */
int talus_mi_plan_get(talus_mi_ctx *ctx, uint8_t *plan, uint32_t capacity, uint32_t *count);
int talus_mi_plan_set(talus_mi_ctx *ctx, const uint8_t *plan, uint32_t count);

struct talus_mi_hex_mesh; /* the hex mesh, defined with the lattices below */
/* collectSurface's hex faces (src/talus/deformable/solver/solver.cc:252-
 * 303) with clusterSurface, for a host hex mesh: each cell face split in two
 * triangles around its lowest node, kept when no other cell has it, turned
 * outward; triangle_tet holds the cell, tets the cell count, surface_nodes
 * the corners of unshared cell faces, crack_faces zero. */
int talus_mi_hex_mesh_surface(const struct talus_mi_hex_mesh *mesh, talus_mi_surface *out);

/* Surface of a context body's current mesh (after its tears), as
 * talus_mi_mesh_surface with the FP64 rest positions. */
int talus_mi_surface_get(talus_mi_ctx *ctx, uint32_t body, talus_mi_surface *out);

/* The context's HIP stream (hipStream_t) for interop. */
void *talus_mi_stream(talus_mi_ctx *ctx);

/* Splits the context's bodies over `streams` engines (1..TALUS_MI_MAX_STREAMS),
 * each with its own HIP stream, stepped concurrently by one host thread each
 * inside talus_mi_step; the state carries over.  No counterpart in the
 * reference: a batch's preconditioner runs level by level, and with two or
 * three groups one group's small top levels overlap another group's work.
 * Bodies go to the group with the fewest tets, largest first.  Reports,
 * events and torn lists are merged in body order; plans are the groups'
 * plans one after another; kernel timings average over every group's
 * launches.  Differentiation (talus_mi_rollout*) needs one stream.
 * Default: 1. */
#define TALUS_MI_MAX_STREAMS 4u
int talus_mi_set_streams(talus_mi_ctx *ctx, uint32_t streams);

/* ---- strainstep hex lattices -------------------------------------------
 * The deformable path World::step runs today: a LatticeBody of hexes solved
 * by solveStrainstep (src/talus/dynamics/strainstep/strainstep.cc:569-842)
 * with one 18-row soft-constraint element per hex
 * (strainstep/element.cc:244-523).  FP64 throughout, like the reference. */

/* Hex PreparedMesh (src/talus/mesh.h:18-44): corner c of a hex sits at
 * offsets (c&1, c>>1&1, c>>2&1) of its cell.  Arrays are caller-owned and
 * copied by talus_mi_lattice_add; the builder allocates them (free with
 * talus_mi_hex_mesh_free). */
typedef struct talus_mi_hex_mesh {
  uint32_t nodes;
  uint32_t hexes;
  float *rest_positions; /* 3*nodes */
  float *node_mass;      /* nodes */
  uint8_t *anchored;     /* nodes */
  uint32_t *hex_nodes;   /* 8*hexes */
  float *volume;         /* hexes */
  uint32_t *hex_material;/* hexes, material lane */
  float voxel_size[3];   /* PreparedMesh::voxelSize: cell size and colours */
} talus_mi_hex_mesh;

/* buildMesh with VolumeElementType::Hex (src/talus/deformable/solver/
 * solver.cc:2185-2316) on a structured lattice: first-touch nodes in a
 * z, y, x cell scan, volume = voxel product, each corner gets rho V / 8 of
 * mass (summed in float), nodes in a support volume anchored. */
int talus_mi_build_hex_lattice(const talus_mi_lattice *lattice,
                               const talus_mi_material *materials, uint32_t lanes,
                               const talus_mi_support *supports, uint32_t nsupports,
                               talus_mi_hex_mesh *out);
void talus_mi_hex_mesh_free(talus_mi_hex_mesh *mesh);

/* makeElement's outputs (StrainElement, strainstep/element.h:40-73) for one
 * hex: host code, no device needed. */
typedef struct talus_mi_hex_element {
  double gradient[24];      /* per corner: rest shape^-T * corner signs / 8 */
  double inverse_mass[8];
  double rest_hourglass[12];
  double bend;
  double size;
  double stiffness[18];
  double rate[18];
  double softness[18];
  double cap[18];
  double inverse[324];      /* 18x18 row-major */
  double volume_stiffness;
  double volume_rate;
  double volume_softness;
  double volume_cap;
  uint32_t volumetric;
  uint32_t reserved;
} talus_mi_hex_element;

/* cornerRest + elementMaterial + makeElement (element.cc:185-190, 199-214,
 * 244-342) for an undamped material, from the eight rest corners (3 each)
 * and the corners' inverse masses; TALUS_MI_ERR_ARGUMENT when the reference
 * builds no element (degenerate rest shape, no shear modulus, no volume or
 * no free corner). */
int talus_mi_hex_element_make(const double *corners, const double *inverse_mass, double volume,
                              float youngs_modulus, float poissons_ratio, double h,
                              talus_mi_hex_element *out);

/* Step parameters: IntegrationParameters (integrationparameters.h:9-33)
 * subset the strainstep path reads. */
/* lattice step flag: time the update and the solve launches with HIP events
 * around each back-to-back run of them (the report's solve_ms / update_ms);
 * otherwise those stay 0 */
#define TALUS_MI_KERNEL_TIMING 16u

typedef struct talus_mi_lattice_params {
  double dt;                  /* default 1/60 */
  double gravity[3];          /* default (0, -9.81, 0) */
  uint32_t softstep_substeps; /* default 4; islandSubsteps' floor */
  uint32_t flags;             /* TALUS_MI_DIRECT_LAUNCH | TALUS_MI_KERNEL_TIMING */
} talus_mi_lattice_params;

typedef struct talus_mi_lattice_report {
  uint32_t lattices;        /* bodies stepped */
  uint32_t substeps;        /* largest over bodies */
  uint32_t elements;        /* hex elements solved */
  uint32_t rebuilds;        /* bodies whose elements were built this step */
  uint64_t element_solves;  /* solveElement calls */
  double device_ms;         /* GPU time of the step's launches */
  uint32_t inverses;        /* distinct 18x18 row-block inverses on the device */
  uint32_t solve_launches;  /* launches of the solve kernel this step */
  uint32_t update_launches; /* launches of the update kernel this step */
  uint32_t refreshed;       /* elements rebuilt for plastic flow (refreshFlowed) */
  double solve_ms;          /* their summed HIP-event time (KERNEL_TIMING) */
  double update_ms;
  uint32_t torn;            /* cells found overstretched at the step's end */
  uint32_t node_copies;     /* nodes split off by the tears they caused */
} talus_mi_lattice_report;

void talus_mi_lattice_default_params(talus_mi_lattice_params *params);

/* addLattice / makeLattice (src/talus/dynamics/bodyset/bodyset.cc:13-86,
 * 120-136): positions rest + offset, velocity given (zero when anchored),
 * cell size = the smallest voxel side.  The strainstep path needs material
 * lanes (strainstep.cc:86-94: lanes past the array read the last one).
 * Lanes with a finite yield stress flow plastically (flowLattice after each
 * step, refreshFlowed before the next).  Lanes with a finite tear strain
 * tear (overstretched / tearCells after each step).  Lanes with a finite
 * damping ratio damp each row toward it at the lattice's lowest angular
 * frequency (cellDamping, strainstep.cc:151-160), found once at the first
 * build; a damped lattice whose banded lowest-mode solve would exceed 4e10
 * operations or 1e8 band entries is refused with TALUS_MI_ERR_ARGUMENT. */
int talus_mi_lattice_add(talus_mi_ctx *ctx, const talus_mi_hex_mesh *mesh,
                         const talus_mi_material *materials, uint32_t lanes,
                         const double *offset, const double *velocity, uint32_t *lattice);
/* LatticeBody::substeps: the body asks for at least this many substeps. */
int talus_mi_lattice_set_substeps(talus_mi_ctx *ctx, uint32_t lattice, uint32_t substeps);
int talus_mi_lattice_state_get(talus_mi_ctx *ctx, uint32_t lattice, double *x, double *v);
int talus_mi_lattice_state_set(talus_mi_ctx *ctx, uint32_t lattice, const double *x,
                               const double *v);
/* The body's elements (hexes that built one, in cell order): cells and the
 * accumulated impulses, 18 rows then the volume row (19 per element).  Either
 * output may be NULL; *count = elements (0 before the first step). */
int talus_mi_lattice_impulses_get(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *cells,
                                  double *impulses, uint32_t capacity, uint32_t *count);
/* LatticeBody::plastic / hardening (plasticflow.cc): per hex and Gauss point
 * (point p at the corner-p signs / sqrt(3)) the inverse plastic deformation P
 * (elastic = F P, 9 per point, row-major) and the accumulated plastic strain
 * (1 per point).  *present = 0 (outputs untouched) until the body first
 * flows.  Setting P makes the next step rebuild the elements of the cells
 * whose P moved (refreshFlowed). */
int talus_mi_lattice_plastic_get(talus_mi_ctx *ctx, uint32_t lattice, double *plastic,
                                 double *hardening, uint32_t *present);
int talus_mi_lattice_plastic_set(talus_mi_ctx *ctx, uint32_t lattice, const double *plastic,
                                 const double *hardening);
/* Node and hex colours (makeLattice, bodyset.cc:30-37): the parity of a
 * node's lattice index per axis; a hex takes its corner 0's colour.  Hexes of
 * one colour share no node, so each colour's elements are solved together. */
int talus_mi_lattice_colours_get(talus_mi_ctx *ctx, uint32_t lattice, uint8_t *node_colour,
                                 uint8_t *hex_colour);
/* The body's current topology: *nodes = its node count (tearing adds node
 * copies), and, each when not NULL, the hex corner nodes (8 per hex), the node
 * masses and per hex its crack-face bits (bit 2 axis + side: the face whose
 * grid neighbour's facing face no longer holds the same nodes; zero before
 * the first tear).  PreparedMesh::hexes / nodeMass / crackFaces after
 * tearCells (src/talus/dynamics/strainstep/tearing.cc:426-484). */
int talus_mi_lattice_topology_get(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *nodes,
                                  uint32_t *hex_nodes, float *node_mass, uint8_t *crack_faces);
/* tearCells (tearing.cc:505-536) at the body's current positions: a crack
 * through each listed cell, the pinched nodes split, node masses recomputed,
 * crack faces marked; *copies = the node copies made.  A torn body's
 * elements are built again at its next step, keeping each cell's impulses. */
int talus_mi_lattice_tear(talus_mi_ctx *ctx, uint32_t lattice, const uint32_t *cells,
                          uint32_t count, uint32_t *copies);
/* cellStretch (tearing.cc:279-281): the largest principal stretch of the
 * cell's eight corner frames and its direction (sign: the largest component
 * positive). */
int talus_mi_lattice_cell_stretch(talus_mi_ctx *ctx, uint32_t lattice, uint32_t cell,
                                  double *stretch, double *direction);
/* StrainLattice::lowest (strainstep.cc:300-331, lowestAngular in
 * lowestmode.cc:31-203): the lowest angular frequency (rad/s) the body's
 * elements are damped for; 0 before a damped body's first step and for an
 * undamped body. */
int talus_mi_lattice_lowest_get(talus_mi_ctx *ctx, uint32_t lattice, double *lowest);
/* PreparedMesh::surfaceNodes, ascending: buildMesh's surface nodes (the
 * corners of every face no other hex shares, solver.cc:2369-2413) with
 * markCracks's additions after tears (crack-face nodes and the copies of
 * surface nodes, tearing.cc:426-484).  *count = the number; the nodes are
 * written when nodes is not NULL and capacity holds them. */
int talus_mi_lattice_surface_get(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *nodes,
                                 uint32_t capacity, uint32_t *count);
/* The hex surface (talus_mi_hex_mesh_surface) of a lattice's current mesh
 * after its tears, with its surface nodes (surfaceNodes) and crack faces
 * (bit 2 axis + side per cell). */
int talus_mi_lattice_surface_mesh(talus_mi_ctx *ctx, uint32_t lattice, talus_mi_surface *out);
/* latticePieces (resolution.cc:198-228): pieces of cells joined by nodes. */
int talus_mi_lattice_pieces(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *pieces);
/* The FP64 stable neo-Hookean hex of tier switching and the energy ledger
 * (src/talus/dynamics/neohookean/neohookean.cc, 8 Gauss points per cell;
 * moduli per cell from its lane, moduliOf 82-90) at positions x of the
 * lattice (3 per node; NULL: its current positions).  cell_energy: cellEnergy
 * with corner 0's displacement taken out of every corner (as latticeEnergy
 * does, ledger.cc:29-47), one per cell; node_gradient (3 per node) and
 * node_hessian (9 per node, row-major): nodeTerms (153-196) summed over the
 * node's cells.  Any output may be NULL. */
int talus_mi_lattice_snh_terms(talus_mi_ctx *ctx, uint32_t lattice, const double *x, double *cell_energy,
                               double *node_gradient, double *node_hessian);
/* Energy (ledger.h:9-15). */
typedef struct talus_mi_energy {
  double kinetic;
  double gravity;
  double elastic;
} talus_mi_energy;
/* latticeEnergy (src/talus/dynamics/ledger/ledger.cc:9-52) at the lattice's
 * current state: kinetic of free nodes, -m g.x of every node, the cells'
 * neo-Hookean energies. */
int talus_mi_lattice_energy(talus_mi_ctx *ctx, uint32_t lattice, const double *gravity,
                            talus_mi_energy *out);
/* RestBody (src/talus/dynamics/screen/screen.h:17-25): total mass, mass
 * centre, inertia about it (row-major), principal axes (columns of a proper
 * rotation) with their moments ascending, and the smallest rest extent. */
typedef struct talus_mi_rest_body {
  double total;
  double centre[3];
  double inertia[9];
  double axes[9];
  double moments[3];
  double thickness;
} talus_mi_rest_body;
/* restBodyOf (screen.cc:26-83) of the lattice's rest mesh. */
int talus_mi_lattice_rest_body(talus_mi_ctx *ctx, uint32_t lattice, talus_mi_rest_body *out);
/* factorScreen + screenLoad (screen.cc:85-254): the displacement (3 per
 * node) of the lattice's rest stiffness under node forces (3 per node), the
 * lattice's anchored nodes and `held` fixed (per-DOF shift 1e6 x the largest
 * stiffness diagonal; free nodes 1e-9 x largest / heaviest x node mass), the
 * load relieved of its rigid acceleration and the rigid part removed from
 * the displacement when no node is held, and dent = max |u| / thickness.
 * Solved on the device by block-Jacobi PCG in FP64 to 1e-12 relative
 * (iterations: the count taken, may be NULL); TALUS_MI_ERR_STATE when it
 * does not converge. */
int talus_mi_lattice_screen(talus_mi_ctx *ctx, uint32_t lattice, const uint32_t *held, uint32_t nheld,
                            const double *forces, double *displacement, double *dent, uint32_t *iterations);
/* A sampled signed distance field (SampledSdf, src/talus/field/sampledsdf/
 * sampledsdf.h:19-32, without its brick summaries): samples x fastest,
 * index x + resolution[0] (y + resolution[1] z); materials may be NULL. */
typedef struct talus_mi_sdf {
  float origin[3];
  float voxel_size[3];
  uint32_t resolution[3];
  float *distances;
  uint8_t *materials;
} talus_mi_sdf;
/* Pose (src/talus/pose.h:9-12): position and rotation quaternion w, x, y, z. */
typedef struct talus_mi_pose {
  float position[3];
  float rotation[4];
} talus_mi_pose;
/* SampleBox (sampledsdf.h:34-37). */
typedef struct talus_mi_sample_box {
  uint32_t minimum[3];
  uint32_t extent[3];
} talus_mi_sample_box;
/* bakeDeformedField (src/talus/deformable/deformedfield/deformedfield.cc:
 * 15-102) of the lattice at its current positions: every sample of the
 * rest field's lattice, placed in the world by the pose, finds the lattice
 * cell holding it (probeDeformed, contact/contact.cc:1287-1425: Newton
 * inversion of the deformed trilinear cell, crack faces capping the
 * distance) and reads the rest field at the matching rest point; samples no
 * cell reaches keep the rest value (or max(rest, pad) with
 * carve_beyond_reach).  Samples that differ from `current` by more than
 * 1e-7 are written into current->distances (in place, on the host) and
 * `changed` receives their box (extent 0 when none); the caller's field
 * refreshes its brick summaries over that box (field::write).  current must
 * share rest's lattice (origin, voxel size, resolution), else
 * TALUS_MI_ERR_ARGUMENT.  One device thread per sample. */
int talus_mi_lattice_bake_field(talus_mi_ctx *ctx, uint32_t lattice, const talus_mi_sdf *rest,
                                const talus_mi_pose *pose, int carve_beyond_reach, talus_mi_sdf *current,
                                talus_mi_sample_box *changed);
/* One solveStrainstep over every lattice of the context, each its own
 * island (no contacts, joints or rigid partners): refreshFlowed of the cells
 * the last step's flow moved, per body islandSubsteps (strainstep.cc:551-567)
 * substeps of integrate, element update and warm-start, 4 Gauss-Seidel passes
 * of solveElement, advance, then flowLattice, and overstretched / tearCells
 * for lattices with a tearing lane (a torn body's elements are rebuilt at its
 * next step with each cell's impulses kept).  The passes
 * sweep the hexes colour by colour (all bodies' hexes of one colour in one
 * launch), where the reference sweeps them in cell order. */
int talus_mi_lattice_step(talus_mi_ctx *ctx, const talus_mi_lattice_params *params,
                          talus_mi_lattice_report *report);

/* ---- contact generation against rigid partners' sampled fields ---- */

/* Copies a sampled field (a rigid partner's shape, ShapeProbe::field,
 * src/talus/deformable/contact/contact.h:19-26) into device memory, where it
 * stays for every query until the context ends; materials NULL = lane 0
 * everywhere.  Its probeLipschitz (sampledsdf.cc:424-430) is computed over
 * every cell, as finishBricks (sampledsdf.cc:584-606) leaves it.  *id
 * names it in talus_mi_shape. */
int talus_mi_field_add(talus_mi_ctx *ctx, const talus_mi_sdf *field, uint32_t *id);
/* Replaces field id's lattice and samples (after the caller rewrote them). */
int talus_mi_field_set(talus_mi_ctx *ctx, uint32_t id, const talus_mi_sdf *field);
/* probeLipschitz of field id: sqrt(g^2 + 1), g the largest cellLipschitz. */
int talus_mi_field_lipschitz(talus_mi_ctx *ctx, uint32_t id, float *lipschitz);

/* ShapeProbe (contact.h:19-26): a rigid partner placed by its centre of
 * mass and rotation (w, x, y, z); a world point w maps to the field point
 * conj(rotation) (w - center_of_mass) + field_offset. */
typedef struct talus_mi_shape {
  uint32_t field; /* talus_mi_field_add id */
  float center_of_mass[3];
  float rotation[4];
  float field_offset[3];
  float linear_velocity[3];
  float angular_velocity[3];
} talus_mi_shape;

/* The arguments of findContacts (contact.h:85-94) besides the mesh, state
 * and materials. */
typedef struct talus_mi_contact_query {
  float dt;
  float gravity[3];
  float bias;          /* ContactGaps::bias */
  float threshold;     /* ContactGaps::threshold */
  float voxel_size[3]; /* PreparedMesh::voxelSize: duplicates lie within 1e-2 of its largest side */
  int32_t cull;
} talus_mi_contact_query;

/* SurfaceContact (contact.h:64-74). */
typedef struct talus_mi_surface_contact {
  uint32_t nodes[3];
  float weights[3];
  int32_t partner;
  float normal[3];
  float local_point[3];
  float gap;
  float bias;
  float friction;
  uint32_t triangle;
} talus_mi_surface_contact;

/* findContacts (src/talus/deformable/contact/contact.cc:599-880) of a tet
 * body's boundary surface (talus_mi_surface_get: its triangles, lanes and
 * clusters) at the body's state (positions and velocities rounded to float)
 * against `count` shapes: for each shape in order, each surface triangle in
 * order whose swept box meets the shape's and that culling (cull != 0:
 * clusterFar per 32-triangle cluster, farFromShape per triangle, with the
 * field's Lipschitz bound) does not rule out, the closest point of the
 * triangle to the shape (seven starts, projected descent) and, when farther
 * than threshold + skin, the advance to impact within dt; a contact within
 * 1e-2 of the largest voxel side of an earlier contact of the same shape is
 * dropped.  thickness (one per surface triangle: surfaceThickness,
 * talus_mi_surface_thickness) may be NULL (zero skin).  Friction is
 * sqrt(f(triangle lane) f(field material at the contact)), f a lane of the
 * body's materials (0.5 beyond them).  Writes min(found, capacity)
 * contacts; *found = how many there are.  Runs on the body's stream after
 * its pending steps. */
int talus_mi_find_contacts(talus_mi_ctx *ctx, uint32_t body, const talus_mi_shape *shapes, uint32_t count,
                           const talus_mi_contact_query *query, const float *thickness,
                           talus_mi_surface_contact *out, uint32_t capacity, uint32_t *found);

#ifdef __cplusplus
}
#endif

#endif
