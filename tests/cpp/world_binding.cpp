// A C++ host binding libtalus_mi through include/talus_mi.hpp the way a
// talus::World would: build a tet cantilever from a lattice, add it with
// PhysicalMaterial-ordered lanes, step it, read the state.  Prints one line
// per check; with no GPU the context creation reports the device error and
// the host-only checks still run.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "talus_mi.hpp"

int main() {
  talus_mi_material jelly{};
  jelly.friction = 0.5f;
  jelly.density = 1000.0f;
  jelly.strength = std::numeric_limits<float>::infinity();
  jelly.youngs_modulus = 1.0e6f;
  jelly.poissons_ratio = 0.3f;
  jelly.yield_stress = std::numeric_limits<float>::infinity();
  jelly.elastic_model = TALUS_MI_NEO_HOOKEAN;
  jelly.damping_ratio = std::numeric_limits<float>::infinity();
  jelly.tear_strain = std::numeric_limits<float>::infinity();
  jelly.plastic_max = std::numeric_limits<float>::infinity();
  jelly.crush_stress = std::numeric_limits<float>::infinity();
  jelly.crush_limit = 1.0f;
  const std::vector<talus_mi_material> lanes{jelly};

  talus_mi_lattice lattice{};
  lattice.resolution[0] = 11;
  lattice.resolution[1] = 5;
  lattice.resolution[2] = 5;
  lattice.voxel_size[0] = 0.08f;
  lattice.voxel_size[1] = 0.05f;
  lattice.voxel_size[2] = 0.05f;
  talus_mi_support pin{};
  const float affine[12] = {1, 0, 0, 0, 0, 1, 0, -0.1f, 0, 0, 1, -0.1f};
  std::memcpy(pin.world_to_support, affine, sizeof affine);
  pin.half_extents[0] = 0.004f;
  pin.half_extents[1] = 1.0f;
  pin.half_extents[2] = 1.0f;
  talus_mi::Mesh mesh = talus_mi::build_lattice(lattice, lanes, {pin});
  double mass = 0.0;
  for (float m : mesh.node_mass) mass += m;
  std::printf("mesh %zu nodes %zu tets mass %.6f\n", mesh.anchored.size(), mesh.tet_material.size(), mass);
  mesh.node_mass.clear();  // masses from rho V / 4 in FP64

  try {
    talus_mi::Context ctx(0, TALUS_MI_FP64);
    const uint32_t body = ctx.add_body(mesh, lanes);
    talus_mi_step_params params = talus_mi::default_params();
    talus_mi_report report{};
    for (int step = 0; step < 10; ++step) report = ctx.step(params);
    const talus_mi::State s = ctx.state(body);
    double sag = 0.0;
    for (size_t n = 0; n < s.x.size() / 3; ++n) sag = std::fmin(sag, s.x[3 * n + 1] - mesh.rest[3 * n + 1]);
    std::printf("stepped 10 steps nonfinite %u sag %.6e\n", report.non_finite, sag);
    // a ground plane 30 cm below the beam as a rigid partner's sampled field
    std::vector<float> ground(9 * 5 * 5);
    for (uint32_t z = 0; z < 5; ++z)
      for (uint32_t y = 0; y < 5; ++y)
        for (uint32_t x = 0; x < 9; ++x) ground[x + 9 * (y + 5 * z)] = -0.5f + 0.25f * float(y) + 0.3f;
    talus_mi_sdf field{{-0.5f, -0.5f, -0.5f}, {0.25f, 0.25f, 0.25f}, {9, 5, 5}, ground.data(), nullptr};
    talus_mi_shape shape{};
    shape.field = ctx.field_add(field);
    shape.rotation[0] = 1.0f;
    talus_mi_contact_query query{};
    query.dt = params.dt;
    query.threshold = 0.4f;
    query.voxel_size[0] = query.voxel_size[1] = query.voxel_size[2] = 0.05f;
    query.cull = 1;
    const std::vector<talus_mi_surface_contact> contacts = ctx.find_contacts(body, {shape}, query);
    float lowest = 1.0f;
    for (const talus_mi_surface_contact &c : contacts) lowest = std::fmin(lowest, c.gap);
    double bottom = 1.0;
    for (size_t n = 0; n < s.x.size() / 3; ++n) bottom = std::fmin(bottom, s.x[3 * n + 1]);
    std::printf("contacts %zu lowest gap %.6e lowest node %.6e\n", contacts.size(), lowest, bottom + 0.3);
  } catch (const talus_mi::Error &e) {
    std::printf("device unavailable code %d: %s\n", e.code, e.what());
  }
  return 0;
}
