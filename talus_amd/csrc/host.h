// host.h -- host-side types of libtalus_mi (not part of the C ABI).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/talus_mi.h"

namespace tmi {

// Records the calling thread's error message and returns `code`.
int fail(int code, const std::string &message);

talus_mi_material default_material();
const talus_mi_material &lane_material(const talus_mi_material *materials, uint32_t lanes,
                                       uint32_t lane);
void precise_lame(float youngs, float poisson, double *lambda, double *mu);
bool rest_frame(const double *rest, const uint32_t *tet, double *dminv, double *volume);
int build_lattice(const talus_mi_lattice *lattice, const talus_mi_material *materials,
                  uint32_t lanes, const talus_mi_support *supports, uint32_t nsupports,
                  talus_mi_mesh *out);
void mesh_free(talus_mi_mesh *mesh);

// One deformable tet body: the FP64 mesh data the device arrays are built
// from, plus a host mirror of the state used across rebuilds.
struct Body {
  uint32_t nodes = 0, tets = 0;
  std::vector<double> rest;       // 3*nodes
  std::vector<double> mass;       // nodes
  std::vector<uint8_t> anchored;  // nodes
  std::vector<uint32_t> tet;      // 4*tets, body-local node indices
  std::vector<double> dminv;      // 9*tets row-major
  std::vector<double> volume;     // tets
  std::vector<uint32_t> lane;     // tets
  std::vector<talus_mi_material> materials;
  // per-tet material data; lame_set keeps lambda and mu as given
  // (talus_mi_lame_set) instead of deriving them from the lanes
  std::vector<double> lambda, mu, yield, hardening, plastic_max, tear;
  std::vector<double> tau;  // stiffnessDamping per tet (0 undamped)
  std::vector<double> rho;  // density when masses follow rho V / 4, else 0
  bool lame_set = false;
  std::vector<uint8_t> flags;
  // state mirror (valid when the device copy is not)
  std::vector<double> x, v, a;
  std::vector<double> R;        // 9*tets
  std::vector<double> W;        // 9*tets, neo-Hookean volume frames
  std::vector<double> plastic;  // 10*tets
  bool density_mass = true;     // masses follow rho V / 4 (recomputed on tears)
};

// Fills per-tet material data and FP64 rest frames; returns a status.
int prepare_body(Body &body);

// Boundary surface of a tet mesh (surface.cpp): outward triangles in
// ascending order of their sorted node triple with their tet and lane,
// unique sorted edges, surface nodes, and per tet the faces that are crack
// faces (bit f for the corner triple f of {0,1,2}, {0,1,3}, {0,2,3},
// {1,2,3}).
struct SurfaceMesh {
  std::vector<uint32_t> triangles;  // 3 per triangle
  std::vector<uint32_t> triangle_tet, triangle_lane;
  std::vector<uint32_t> edges;  // 2 per edge
  std::vector<uint32_t> nodes;
  std::vector<uint8_t> crack;   // per tet
  // clusterSurface (contact.cc:540-597): triangles in Morton order of their
  // rest centres, SURFACE_CLUSTER = 32 to a cluster; starts end with the
  // triangle count
  std::vector<uint32_t> cluster_faces, cluster_starts;
};
void mesh_surface(uint32_t nodes, uint32_t tets, const uint32_t *tet, const uint32_t *lane,
                  const double *rest, SurfaceMesh &out);
// collectSurface's hex faces (solver.cc:252-303): each cell face split in
// two triangles around its lowest node, a face kept when no other cell has
// it, turned outward by the cell's opposite corner; triangles in ascending
// order of their sorted node triple, triangle_tet the cell, edges the
// triangles' sorted pairs without repeats, clusters as for tets.  Surface
// nodes and crack faces are the caller's (LatticeBody).
void hex_surface(uint32_t nodes, uint32_t hexes, const uint32_t *hex, const uint32_t *lane, const double *rest,
                 SurfaceMesh &out);
// field::distance (sampledsdf.cc:377-385) in float: the distance to the
// lattice's box outside it, else the trilinear sample
float sdf_distance(const talus_mi_sdf &field, const float *point);
// surfaceThickness of collectSurface (solver.cc:342-351): minus the mean
// field distance at each triangle's three rest positions
void surface_thickness(const talus_mi_sdf &field, const double *rest, uint32_t triangles,
                       const uint32_t *triangle_nodes, float *thickness);
// clusterSurface of outward triangles (3 node indices each) over rest
// positions that hold float samples
void cluster_surface(const std::vector<uint32_t> &triangles, const double *rest, std::vector<uint32_t> &faces,
                     std::vector<uint32_t> &starts);

// Tear topology update (tear.cpp): splits nodes around the torn tets
// (ascending body-local indices) at the body's positions b.x; new nodes are
// appended and origin[i] is the node that node (old count + i) came from.
// Returns the number of new nodes.  Masses are not recomputed here.
uint32_t tear_topology(Body &body, const std::vector<uint32_t> &torn, std::vector<uint32_t> &origin);

}  // namespace tmi
