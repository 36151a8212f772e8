// talus_mi.hpp -- header-only C++17 wrapper over the libtalus_mi C ABI
// (include/talus_mi.h): RAII context, exceptions instead of status codes,
// std::vector in and out.  It adds no behaviour; every call maps to one C
// entry point, so a C++ host (a talus::World) binds the library with this
// header and -ltalus_mi.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "talus_mi.h"

namespace talus_mi {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

inline void check(int status) {
  if (status != TALUS_MI_OK) {
    const char *message = talus_mi_last_error();
    throw Error(status, message ? message : "talus_mi error");
  }
}

inline talus_mi_step_params default_params() {
  talus_mi_step_params p;
  talus_mi_default_params(&p);
  return p;
}

// Host tet mesh in the C ABI's layout: rest positions xyz per node, anchored
// flag per node, 4 node indices and a material lane per tet; node_mass empty
// means rho V / 4 per incident tet.
struct Mesh {
  std::vector<float> rest;
  std::vector<uint8_t> anchored;
  std::vector<uint32_t> tet_nodes;
  std::vector<uint32_t> tet_material;
  std::vector<float> node_mass;

  talus_mi_mesh view() {
    talus_mi_mesh m{};
    m.nodes = uint32_t(anchored.size());
    m.tets = uint32_t(tet_material.size());
    m.rest_positions = rest.data();
    m.anchored = anchored.data();
    m.tet_nodes = tet_nodes.data();
    m.tet_material = tet_material.data();
    m.node_mass = node_mass.empty() ? nullptr : node_mass.data();
    return m;
  }
};

struct State {
  std::vector<double> x, v, a;  // 3 per node
};

class Context {
 public:
  explicit Context(int device = 0, uint32_t precision = TALUS_MI_MIXED) {
    check(talus_mi_create(&ctx_, device, precision));
  }
  ~Context() {
    if (ctx_) talus_mi_destroy(ctx_);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  Context(Context &&o) noexcept : ctx_(std::exchange(o.ctx_, nullptr)) {}
  Context &operator=(Context &&o) noexcept {
    if (this != &o) {
      if (ctx_) talus_mi_destroy(ctx_);
      ctx_ = std::exchange(o.ctx_, nullptr);
    }
    return *this;
  }

  talus_mi_ctx *get() const { return ctx_; }

  uint32_t add_body(Mesh &mesh, const std::vector<talus_mi_material> &lanes) {
    talus_mi_mesh m = mesh.view();
    uint32_t id = 0;
    check(talus_mi_body_add(ctx_, &m, lanes.data(), uint32_t(lanes.size()), &id));
    return id;
  }

  std::pair<uint32_t, uint32_t> size(uint32_t body) const {
    uint32_t nodes = 0, tets = 0;
    check(talus_mi_body_info(ctx_, body, &nodes, &tets));
    return {nodes, tets};
  }

  State state(uint32_t body) const {
    const uint32_t n = size(body).first;
    State s{std::vector<double>(3 * size_t(n)), std::vector<double>(3 * size_t(n)),
            std::vector<double>(3 * size_t(n))};
    check(talus_mi_state_get(ctx_, body, s.x.data(), s.v.data(), s.a.data()));
    return s;
  }

  void set_state(uint32_t body, const State &s) {
    check(talus_mi_state_set(ctx_, body, s.x.empty() ? nullptr : s.x.data(),
                             s.v.empty() ? nullptr : s.v.data(), s.a.empty() ? nullptr : s.a.data()));
  }

  talus_mi_report step(const talus_mi_step_params &params) {
    talus_mi_report report;
    check(talus_mi_step(ctx_, &params, &report));
    return report;
  }

  std::vector<talus_mi_event> events(uint32_t capacity = 1u << 16) {
    std::vector<talus_mi_event> out(capacity);
    uint32_t n = 0;
    check(talus_mi_events(ctx_, out.data(), capacity, &n));
    out.resize(n);
    return out;
  }

  // connectivity, rest positions, masses and anchors after tears
  void topology(uint32_t body, std::vector<uint32_t> &tet_nodes, std::vector<double> &rest,
                std::vector<double> &mass, std::vector<uint8_t> &anchored) const {
    const auto [n, t] = size(body);
    tet_nodes.resize(4 * size_t(t));
    rest.resize(3 * size_t(n));
    mass.resize(n);
    anchored.resize(n);
    check(talus_mi_topology_get(ctx_, body, tet_nodes.data(), rest.data(), mass.data(),
                                anchored.data()));
  }

  // a rigid partner's sampled field, resident on the device from now on
  uint32_t field_add(const talus_mi_sdf &field) {
    uint32_t id = 0;
    check(talus_mi_field_add(ctx_, &field, &id));
    return id;
  }
  void field_set(uint32_t id, const talus_mi_sdf &field) { check(talus_mi_field_set(ctx_, id, &field)); }

  // findContacts of a body's surface against the shapes; thickness (one per
  // surface triangle) may be empty
  std::vector<talus_mi_surface_contact> find_contacts(uint32_t body, const std::vector<talus_mi_shape> &shapes,
                                                      const talus_mi_contact_query &query,
                                                      const std::vector<float> &thickness = {}) {
    std::vector<talus_mi_surface_contact> out(256);
    uint32_t found = 0;
    for (;;) {
      check(talus_mi_find_contacts(ctx_, body, shapes.data(), uint32_t(shapes.size()), &query,
                                   thickness.empty() ? nullptr : thickness.data(), out.data(),
                                   uint32_t(out.size()), &found));
      if (found <= out.size()) break;
      out.resize(found);
    }
    out.resize(found);
    return out;
  }

 private:
  talus_mi_ctx *ctx_ = nullptr;
};

// Host lattice -> tet mesh (Kuhn split, first-touch node order); the result
// is copied into a Mesh and the library's allocation released.
inline Mesh build_lattice(const talus_mi_lattice &lattice, const std::vector<talus_mi_material> &lanes,
                          const std::vector<talus_mi_support> &supports) {
  talus_mi_mesh out{};
  check(talus_mi_build_lattice(&lattice, lanes.data(), uint32_t(lanes.size()), supports.data(),
                               uint32_t(supports.size()), &out));
  Mesh m;
  m.rest.assign(out.rest_positions, out.rest_positions + 3 * size_t(out.nodes));
  m.anchored.assign(out.anchored, out.anchored + out.nodes);
  m.tet_nodes.assign(out.tet_nodes, out.tet_nodes + 4 * size_t(out.tets));
  m.tet_material.assign(out.tet_material, out.tet_material + out.tets);
  m.node_mass.assign(out.node_mass, out.node_mass + out.nodes);
  talus_mi_mesh_free(&out);
  return m;
}

}  // namespace talus_mi
