// contact_math.h -- the float geometry of contact generation
// (src/talus/deformable/contact/contact.cc:10-536 and the float probe of
// src/talus/field/sampledsdf/sampledsdf.cc:18-93, 234-305, 387-477), in the
// reference's order of operations.  Host and device code both compile it;
// contact.hip builds it with contraction off and correctly rounded division
// and square root, so that each operation rounds as the reference's does.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "../../include/talus_mi.h"

#if defined(__HIPCC__)
#define TMI_HD __host__ __device__
#else
#define TMI_HD
#endif

namespace tmi {

// a sampled field's lattice and its samples (device memory in the kernels)
struct FieldView {
  float origin[3], voxel[3];
  uint32_t res[3];
  const float *dist;
  const uint8_t *mat;
};

namespace contact_math {

constexpr uint32_t SURFACE_CLUSTER = 32u;   // contact.h:81
constexpr uint32_t ADVANCE_ITERATIONS = 4u;  // contact.cc:73-77
constexpr uint32_t DESCENT_ITERATIONS = 8u;
constexpr uint32_t BACKTRACKS = 4u;
constexpr float IMPACT_RELATIVE_ERROR = 0.05f;
constexpr float CULL_TOLERANCE = 1.0e-5f;
constexpr float GRADIENT_FLOOR = 1e-6f;      // sampledsdf.cc:7
constexpr float DEFAULT_FRICTION = 0.5f;     // PhysicalMaterial::friction

struct F3 {
  float x, y, z;
};
struct Q4 {
  float w, x, y, z;
};

TMI_HD inline F3 mk(float x, float y, float z) { return F3{x, y, z}; }
TMI_HD inline F3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
TMI_HD inline F3 operator+(F3 a, F3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
TMI_HD inline F3 operator-(F3 a, F3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
TMI_HD inline F3 operator-(F3 a) { return mk(-a.x, -a.y, -a.z); }
TMI_HD inline F3 operator*(F3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
TMI_HD inline F3 operator/(F3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
TMI_HD inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TMI_HD inline float len(F3 a) { return sqrtf(dot(a, a)); }
TMI_HD inline F3 cross(F3 a, F3 b) {
  return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
TMI_HD inline float mn(float a, float b) { return b < a ? b : a; }
TMI_HD inline float mx(float a, float b) { return a < b ? b : a; }
TMI_HD inline F3 vmin(F3 a, F3 b) { return mk(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
TMI_HD inline F3 vmax(F3 a, F3 b) { return mk(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }

TMI_HD inline Q4 qmul(Q4 p, Q4 q) {
  return Q4{p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z, p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
            p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z, p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x};
}
TMI_HD inline Q4 qconj(Q4 q) { return Q4{q.w, -q.x, -q.y, -q.z}; }
// q v q*: v + 2 (w (u x v) + u x (u x v)), u the vector part
TMI_HD inline F3 rotate(Q4 q, F3 v) {
  const F3 u = mk(q.x, q.y, q.z);
  const F3 uv = cross(u, v);
  const F3 uuv = cross(u, uv);
  return v + (uv * q.w + uuv) * 2.0f;
}

// ------------------------------------------------------------ the field

struct DevShape {
  FieldView f;
  float center[3], rot[4], offset[3], lin[3], ang[3];
  float lipschitz;
};

TMI_HD inline F3 field_low(const FieldView &f) { return ld3(f.origin); }
TMI_HD inline F3 field_high(const FieldView &f) {
  return mk(f.origin[0] + f.voxel[0] * float(f.res[0] - 1u), f.origin[1] + f.voxel[1] * float(f.res[1] - 1u),
            f.origin[2] + f.voxel[2] * float(f.res[2] - 1u));
}

struct Cell {
  uint32_t base[3];
  float t[3];
};

TMI_HD inline Cell locate(const FieldView &f, F3 p) {
  Cell c;
  const float lattice[3] = {(p.x - f.origin[0]) / f.voxel[0], (p.y - f.origin[1]) / f.voxel[1],
                            (p.z - f.origin[2]) / f.voxel[2]};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const uint32_t b = uint32_t(floorf(lattice[k]));
    c.base[k] = b < f.res[k] - 2u ? b : f.res[k] - 2u;
    c.t[k] = lattice[k] - float(c.base[k]);
  }
  return c;
}

TMI_HD inline void corners(const FieldView &f, const Cell &c, float *v) {
  const size_t sy = f.res[0], sz = size_t(f.res[0]) * f.res[1];
  const size_t base = c.base[0] + sy * c.base[1] + sz * c.base[2];
  v[0] = f.dist[base];
  v[1] = f.dist[base + 1];
  v[2] = f.dist[base + sy];
  v[3] = f.dist[base + sy + 1];
  v[4] = f.dist[base + sz];
  v[5] = f.dist[base + sz + 1];
  v[6] = f.dist[base + sz + sy];
  v[7] = f.dist[base + sz + sy + 1];
}

TMI_HD inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

TMI_HD inline float interpolate(const float *v, const Cell &c) {
  const float x00 = mixf(v[0], v[1], c.t[0]);
  const float x10 = mixf(v[2], v[3], c.t[0]);
  const float x01 = mixf(v[4], v[5], c.t[0]);
  const float x11 = mixf(v[6], v[7], c.t[0]);
  const float y0 = mixf(x00, x10, c.t[1]);
  const float y1 = mixf(x01, x11, c.t[1]);
  return mixf(y0, y1, c.t[2]);
}

TMI_HD inline F3 slope(const FieldView &f, const float *c, const Cell &cell) {
  const float u = cell.t[0], v = cell.t[1], w = cell.t[2];
  const float dx = (c[1] - c[0]) * (1.0f - v) * (1.0f - w) + (c[3] - c[2]) * v * (1.0f - w) +
                   (c[5] - c[4]) * (1.0f - v) * w + (c[7] - c[6]) * v * w;
  const float dy = (c[2] - c[0]) * (1.0f - u) * (1.0f - w) + (c[3] - c[1]) * u * (1.0f - w) +
                   (c[6] - c[4]) * (1.0f - u) * w + (c[7] - c[5]) * u * w;
  const float dz = (c[4] - c[0]) * (1.0f - u) * (1.0f - v) + (c[5] - c[1]) * u * (1.0f - v) +
                   (c[6] - c[2]) * (1.0f - u) * v + (c[7] - c[3]) * u * v;
  return mk(dx / f.voxel[0], dy / f.voxel[1], dz / f.voxel[2]);
}

struct Probe {
  float distance;
  F3 normal;
};

// probe<float> (sampledsdf.cc:432-477)
TMI_HD inline Probe field_probe(const FieldView &f, F3 p) {
  const F3 clamped = vmin(vmax(p, field_low(f)), field_high(f));
  const F3 gap = p - clamped;
  const float outside = len(gap);
  const Cell cell = locate(f, clamped);
  float c[8];
  corners(f, cell, c);
  const float inner = interpolate(c, cell);
  Probe r;
  r.distance = outside <= 0.0f ? inner : inner + outside;
  const F3 derivative = slope(f, c, cell);
  const float steepness = len(derivative);
  const bool usable = steepness > GRADIENT_FLOOR;
  if (outside <= 0.0f) {
    r.normal = usable ? derivative / steepness : mk(0.0f, 1.0f, 0.0f);
    return r;
  }
  const F3 blended = (usable ? derivative / steepness * fabsf(inner) : mk(0.0f, 0.0f, 0.0f)) + gap;
  const float reach = len(blended);
  r.normal = reach > GRADIENT_FLOOR ? blended / reach : gap / outside;
  return r;
}

// probeDistance (sampledsdf.cc:417-422)
TMI_HD inline float probe_distance(const FieldView &f, F3 p) {
  const F3 clamped = vmin(vmax(p, field_low(f)), field_high(f));
  const float outside = len(p - clamped);
  const Cell cell = locate(f, clamped);
  float c[8];
  corners(f, cell, c);
  const float inner = interpolate(c, cell);
  return outside <= 0.0f ? inner : inner + outside;
}

// material (sampledsdf.cc:387-402); a field without materials reads lane 0
TMI_HD inline uint32_t field_material(const FieldView &f, F3 p) {
  const F3 clamped = vmin(vmax(p, field_low(f)), field_high(f));
  if (len(p - clamped) > 0.0f || !f.mat) return 0u;
  const Cell c = locate(f, p);
  const uint32_t x = c.base[0] + (c.t[0] < 0.5f ? 0u : 1u), y = c.base[1] + (c.t[1] < 0.5f ? 0u : 1u),
                 z = c.base[2] + (c.t[2] < 0.5f ? 0u : 1u);
  return f.mat[x + size_t(f.res[0]) * (y + size_t(f.res[1]) * z)];
}

// ------------------------------------------------------------ the shape

struct ShapeAt {
  F3 center;
  Q4 rotation;
};

TMI_HD inline Q4 shape_rotation(const DevShape &s) { return Q4{s.rot[0], s.rot[1], s.rot[2], s.rot[3]}; }

// advanceShape (contact.cc:10-23)
TMI_HD inline ShapeAt advance_shape(const DevShape &s, float elapsed) {
  ShapeAt at;
  at.center = ld3(s.center) + ld3(s.lin) * elapsed;
  const F3 w = ld3(s.ang);
  const float spin = len(w);
  if (spin > 0.0f) {
    const F3 axis = w / spin;
    const float half = spin * elapsed * 0.5f;
    const float sn = float(sin(double(half)));
    const Q4 turn{float(cos(double(half))), axis.x * sn, axis.y * sn, axis.z * sn};
    at.rotation = qmul(turn, shape_rotation(s));
  } else {
    at.rotation = shape_rotation(s);
  }
  return at;
}

TMI_HD inline F3 field_point_of(const DevShape &s, const ShapeAt &at, F3 world) {
  return rotate(qconj(at.rotation), world - at.center) + ld3(s.offset);
}

TMI_HD inline Probe probe_shape(const DevShape &s, const ShapeAt &at, F3 world) {
  const Probe probed = field_probe(s.f, field_point_of(s, at, world));
  const F3 outward = rotate(at.rotation, probed.normal);
  const float l = len(outward);
  return Probe{probed.distance, l > 0.0f ? outward / l : mk(0.0f, 1.0f, 0.0f)};
}

// ---------------------------------------------------------- the triangle

struct Tri {
  F3 c[3];
};

// barycentricOf (contact.cc:87-156)
TMI_HD inline F3 barycentric(const Tri &t, F3 p) {
  const F3 a = t.c[0], b = t.c[1], c = t.c[2];
  const F3 ab = b - a, ac = c - a, ap = p - a;
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return mk(1.0f, 0.0f, 0.0f);
  const F3 bp = p - b;
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return mk(0.0f, 1.0f, 0.0f);
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    return mk(1.0f - v, v, 0.0f);
  }
  const F3 cp = p - c;
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return mk(0.0f, 0.0f, 1.0f);
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    return mk(1.0f - w, 0.0f, w);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return mk(0.0f, 1.0f - w, w);
  }
  const float denominator = 1.0f / (va + vb + vc);
  const float v = vb * denominator, w = vc * denominator;
  return mk(1.0f - v - w, v, w);
}

TMI_HD inline F3 point_at(const Tri &t, F3 w) { return t.c[0] * w.x + t.c[1] * w.y + t.c[2] * w.z; }

struct Closest {
  F3 weights, point;
  Probe probe;
};

// closestOnTriangle (contact.cc:187-297): seven starts, then up to eight
// projected descent steps with four halvings each
TMI_HD inline Closest closest_on(const Tri &t, const DevShape &s, const ShapeAt &at) {
  const float third = 1.0f / 3.0f;
  Closest best;
  best.weights = mk(0.0f, 0.0f, 0.0f);
  best.point = mk(0.0f, 0.0f, 0.0f);
  best.probe = Probe{FLT_MAX, mk(0.0f, 1.0f, 0.0f)};
  for (int i = 0; i < 7; ++i) {
    const F3 w = i == 0   ? mk(1.0f, 0.0f, 0.0f)
                 : i == 1 ? mk(0.0f, 1.0f, 0.0f)
                 : i == 2 ? mk(0.0f, 0.0f, 1.0f)
                 : i == 3 ? mk(0.5f, 0.5f, 0.0f)
                 : i == 4 ? mk(0.0f, 0.5f, 0.5f)
                 : i == 5 ? mk(0.5f, 0.0f, 0.5f)
                          : mk(third, third, third);
    Closest c;
    c.weights = w;
    c.point = point_at(t, w);
    c.probe = probe_shape(s, at, c.point);
    if (c.probe.distance < best.probe.distance) best = c;
  }
  if (!(best.probe.distance < FLT_MAX)) return best;
  const F3 across = cross(t.c[1] - t.c[0], t.c[2] - t.c[0]);
  const float area = len(across);
  if (area <= 0.0f) return best;
  const F3 facing = across / area;
  const float edge = mx(len(t.c[1] - t.c[0]), mx(len(t.c[2] - t.c[1]), len(t.c[0] - t.c[2])));
  for (uint32_t it = 0; it != DESCENT_ITERATIONS; ++it) {
    const F3 g = best.probe.normal;
    const F3 slide = -(g - facing * dot(facing, g));
    const float l = len(slide);
    if (l <= 1.0e-4f) break;
    const F3 direction = slide / l;
    float stride = mn(mx(fabsf(best.probe.distance), 0.05f * edge), 0.5f * edge);
    bool improved = false;
    for (uint32_t a = 0; a != BACKTRACKS; ++a) {
      Closest c;
      c.weights = barycentric(t, best.point + direction * stride);
      c.point = point_at(t, c.weights);
      c.probe = probe_shape(s, at, c.point);
      if (c.probe.distance < best.probe.distance) {
        best = c;
        improved = true;
        break;
      }
      stride *= 0.5f;
    }
    if (!improved) break;
  }
  return best;
}

// ------------------------------------------------------------ the bounds

struct Bounds {
  F3 lo, hi;
};

TMI_HD inline Bounds empty_bounds() {
  return Bounds{mk(FLT_MAX, FLT_MAX, FLT_MAX), mk(-FLT_MAX, -FLT_MAX, -FLT_MAX)};
}
TMI_HD inline void include(Bounds &b, F3 p) {
  b.lo = vmin(b.lo, p);
  b.hi = vmax(b.hi, p);
}
TMI_HD inline bool overlaps(const Bounds &a, const Bounds &b, float margin) {
  return a.lo.x - margin <= b.hi.x && a.lo.y - margin <= b.hi.y && a.lo.z - margin <= b.hi.z &&
         b.lo.x - margin <= a.hi.x && b.lo.y - margin <= a.hi.y && b.lo.z - margin <= a.hi.z;
}

// sweptShapeBounds (contact.cc:314-339)
TMI_HD inline Bounds swept_bounds(const DevShape &s, float dt) {
  const F3 low = field_low(s.f) - ld3(s.offset), high = field_high(s.f) - ld3(s.offset);
  Bounds b = empty_bounds();
  for (int i = 0; i < 2; ++i) {
    const ShapeAt at = advance_shape(s, i == 0 ? 0.0f : dt);
    for (uint32_t c = 0; c != 8u; ++c) {
      const F3 local = mk((c & 1u) ? high.x : low.x, (c & 2u) ? high.y : low.y, (c & 4u) ? high.z : low.z);
      include(b, at.center + rotate(at.rotation, local));
    }
  }
  return b;
}

TMI_HD inline float shape_radius(const DevShape &s) {
  const F3 low = field_low(s.f) - ld3(s.offset), high = field_high(s.f) - ld3(s.offset);
  return len(vmax(mk(fabsf(low.x), fabsf(low.y), fabsf(low.z)), mk(fabsf(high.x), fabsf(high.y), fabsf(high.z))));
}

TMI_HD inline float impact_bound(const DevShape &s, float speed, float radius) {
  return speed + len(ld3(s.lin)) + len(ld3(s.ang)) * radius;
}

// distanceBelow (contact.cc:382-398)
TMI_HD inline float distance_below(const DevShape &s, const ShapeAt &origin, F3 centre, float spread, float lipschitz) {
  const F3 local = field_point_of(s, origin, centre);
  const float sampled = probe_distance(s.f, local);
  const float magnitude = 1.0f + len(centre) + len(origin.center) + len(local) + len(field_low(s.f)) +
                          fabsf(sampled) + spread;
  return sampled - lipschitz * spread - CULL_TOLERANCE * (1.0f + lipschitz) * magnitude;
}

TMI_HD inline bool out_of_reach(const DevShape &s, float lower, float speed, float dt, float bias, float threshold,
                             float radius) {
  if (!(lower > threshold)) return false;
  const float bound = impact_bound(s, speed, radius);
  return bound <= 0.0f || (lower - bias) / bound > dt;
}

struct Moving {
  Tri start;
  F3 velocity[3];
};

TMI_HD inline Tri moving_at(const Moving &m, float elapsed) {
  Tri t = m.start;
  for (int c = 0; c < 3; ++c) t.c[c] = t.c[c] + m.velocity[c] * elapsed;
  return t;
}

TMI_HD inline float moving_speed(const Moving &m) {
  float fastest = 0.0f;
  for (int c = 0; c < 3; ++c) fastest = mx(fastest, len(m.velocity[c]));
  return fastest;
}

// what every triangle's test reads of a shape: its swept box over dt, its
// radius and its pose at the start (sweptShapeBounds, shapeRadius,
// advanceShape(shape, 0))
struct ShapePre {
  Bounds swept;
  float radius;
  ShapeAt origin;
};

TMI_HD inline ShapePre shape_pre(const DevShape &s, float dt) {
  return ShapePre{swept_bounds(s, dt), shape_radius(s), advance_shape(s, 0.0f)};
}

// advanceToImpact (contact.cc:493-536)
TMI_HD inline bool advance_to_impact(const DevShape &s, const Moving &m, float dt, float bias, float radius,
                                  Closest &closest, float &impact_time) {
  const float bound = impact_bound(s, moving_speed(m), radius);
  if (bound <= 0.0f) return false;
  float elapsed = (closest.probe.distance - bias) / bound;
  if (elapsed > dt) return false;
  const float reach = bias * (1.0f + IMPACT_RELATIVE_ERROR);
  for (uint32_t it = 0; it != ADVANCE_ITERATIONS; ++it) {
    const ShapeAt at = advance_shape(s, elapsed);
    closest = closest_on(moving_at(m, elapsed), s, at);
    impact_time = elapsed;
    if (closest.probe.distance <= reach) return true;
    elapsed += (closest.probe.distance - bias) / bound;
    if (elapsed > dt) return false;
  }
  return true;
}

struct Candidate {
  talus_mi_surface_contact c;
  float point[3];  // the mesh point, for the duplicate test
};

// what the face loop of findContacts reads besides the shape
struct FaceQuery {
  const uint32_t *tri, *lane;
  const float *pos, *vel, *friction;  // float state (3 per node), lane frictions
  uint32_t lanes;
  float dt, bias, threshold;
  int cull;
};

// the body of findContacts' face loop (contact.cc:739-869) for one triangle
// and shape up to the duplicate test: false when the triangle makes no
// contact, else the contact and its mesh point in *o
TMI_HD inline bool face_contact(const FaceQuery &q, const DevShape &sh, const ShapePre &pre, int32_t partner,
                                uint32_t face, float thick, Candidate *o) {
  const float sb = q.bias + thick, st = q.threshold + thick;
  Moving mv;
  Bounds reach = empty_bounds();
  uint32_t nodes[3];
  for (int k = 0; k < 3; ++k) {
    nodes[k] = q.tri[3 * size_t(face) + k];
    mv.start.c[k] = ld3(&q.pos[3 * size_t(nodes[k])]);
    mv.velocity[k] = ld3(&q.vel[3 * size_t(nodes[k])]);
    include(reach, mv.start.c[k]);
    include(reach, mv.start.c[k] + mv.velocity[k] * q.dt);
  }
  if (!overlaps(reach, pre.swept, st)) return false;
  const float radius = pre.radius;
  const ShapeAt &origin = pre.origin;
  if (q.cull) {
    const F3 centre = (mv.start.c[0] + mv.start.c[1] + mv.start.c[2]) / 3.0f;
    float spread = 0.0f;
    for (int k = 0; k < 3; ++k) spread = mx(spread, len(centre - mv.start.c[k]));
    if (out_of_reach(sh, distance_below(sh, origin, centre, spread, sh.lipschitz), moving_speed(mv), q.dt, sb, st,
                     radius))
      return false;
  }
  Closest closest = closest_on(mv.start, sh, origin);
  float impact_time = 0.0f;
  if (closest.probe.distance > st && !advance_to_impact(sh, mv, q.dt, sb, radius, closest, impact_time)) return false;
  const ShapeAt impact = advance_shape(sh, impact_time);
  if (closest.probe.distance <= 0.0f) {
    const Tri now = moving_at(mv, impact_time);
    const F3 across = cross(now.c[1] - now.c[0], now.c[2] - now.c[0]);
    const float area = len(across);
    if (area > 0.0f) closest.probe.normal = -across / area;
  }
  const F3 on_shape = closest.point - closest.probe.normal * closest.probe.distance;
  const F3 local = rotate(qconj(impact.rotation), on_shape - impact.center);
  const F3 mesh_point = point_at(mv.start, closest.weights);
  const F3 shape_point = origin.center + rotate(origin.rotation, local);
  for (int k = 0; k < 3; ++k) o->c.nodes[k] = nodes[k];
  o->c.weights[0] = closest.weights.x, o->c.weights[1] = closest.weights.y, o->c.weights[2] = closest.weights.z;
  o->c.partner = partner;
  o->c.normal[0] = closest.probe.normal.x, o->c.normal[1] = closest.probe.normal.y;
  o->c.normal[2] = closest.probe.normal.z;
  o->c.local_point[0] = local.x, o->c.local_point[1] = local.y, o->c.local_point[2] = local.z;
  o->c.bias = sb;
  const uint32_t lane = q.lane[face];
  const uint32_t partner_lane = field_material(sh.f, field_point_of(sh, impact, on_shape));
  const float mine = lane < q.lanes ? q.friction[lane] : DEFAULT_FRICTION;
  const float theirs = partner_lane < q.lanes ? q.friction[partner_lane] : DEFAULT_FRICTION;
  o->c.friction = sqrtf(mine * theirs);
  o->c.gap = dot(closest.probe.normal, mesh_point - shape_point) - sb;
  o->c.triangle = face;
  o->point[0] = mesh_point.x, o->point[1] = mesh_point.y, o->point[2] = mesh_point.z;
  return true;
}

}  // namespace contact_math
}  // namespace tmi
