// contact.h -- contact generation of a tet body's surface against rigid
// partners' sampled fields on the GPU (contact.hip; the C ABI's
// talus_mi_field_* and talus_mi_find_contacts).
//
// Reference (paths relative to the Talus tree):
//   src/talus/deformable/contact/contact.cc:10-880   findContacts and helpers
//   src/talus/field/sampledsdf/sampledsdf.cc:18-93, 234-305, 387-477
//                                                     the float probe,
//                                                     material, Lipschitz bound
//
// Kernels: one thread per node converts the FP64 state to the float state
// the reference reads; one 32-lane half-wavefront per surface cluster (32
// triangles in Morton order) reduces its swept box, centre, spread, speed and
// skin and tests it against every shape; one thread per (shape, triangle),
// in cluster order so that a wavefront probes one neighbourhood of the
// field, finds the triangle's closest point, advances to impact and writes a
// candidate; one workgroup keeps the candidates in face order and drops
// duplicates (the reference's greedy rule, resolved in parallel rounds).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "../../include/talus_mi.h"
#include "contact_math.h"
#include "host.h"

namespace tmi {

// a sampled field resident in device memory with its probeLipschitz
class DeviceField {
 public:
  DeviceField() = default;
  DeviceField(const DeviceField &) = delete;
  DeviceField &operator=(const DeviceField &) = delete;
  ~DeviceField();
  // copies the samples (and materials, zero when absent) and computes the
  // Lipschitz bound on `stream`; waits for both
  void set(const talus_mi_sdf &field, hipStream_t stream);
  FieldView view() const { return view_; }
  float lipschitz() const { return lipschitz_; }

 private:
  FieldView view_{};
  float *dist_ = nullptr;
  uint8_t *mat_ = nullptr;
  size_t capacity_ = 0;
  float lipschitz_ = 0.0f;
};

// one body's surface on the device and the query's work buffers
class ContactFinder {
 public:
  // the body's boundary surface (mesh_surface with its clusters), triangle
  // lanes, anchored nodes and lane frictions
  explicit ContactFinder(const Body &body);
  ContactFinder(const ContactFinder &) = delete;
  ContactFinder &operator=(const ContactFinder &) = delete;
  ~ContactFinder();
  uint32_t triangles() const { return T_; }
  // findContacts at device positions x and velocities v (FP64, 3 per node)
  // against `count` shapes whose fields are `fields`; thickness (one per
  // triangle, host) may be null; writes up to capacity contacts to out and
  // returns how many there are
  uint32_t find(const double *x, const double *v, hipStream_t stream, const talus_mi_shape *shapes,
                const DeviceField *const *fields, uint32_t count, const talus_mi_contact_query &query,
                const float *thickness, talus_mi_surface_contact *out, uint32_t capacity);

 private:
  void reserve(uint32_t shapes, uint32_t capacity);
  uint32_t N_ = 0, T_ = 0, C_ = 0, lanes_ = 0;
  std::vector<void *> owned_;
  uint32_t *tri_ = nullptr, *lane_ = nullptr, *cfaces_ = nullptr;
  uint8_t *anchored_ = nullptr;
  float *friction_ = nullptr, *thickness_ = nullptr, *pos_ = nullptr, *vel_ = nullptr;
  void *clusters_ = nullptr;  // ClusterData per cluster
  uint32_t *count_ = nullptr;
  // per (shape, triangle), grown with the number of shapes
  uint32_t shapes_cap_ = 0, out_cap_ = 0;
  std::vector<void *> per_shape_;
  contact_math::DevShape *dshapes_ = nullptr;
  contact_math::ShapePre *pre_ = nullptr;
  contact_math::Candidate *cand_ = nullptr;
  uint8_t *skip_ = nullptr, *flag_ = nullptr, *state_ = nullptr;
  uint32_t *order_ = nullptr, *nbr_ = nullptr, *nbn_ = nullptr, *kount_ = nullptr;
  float *pts_ = nullptr;
  void *out_ = nullptr;
};

}  // namespace tmi
