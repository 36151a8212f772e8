/* contact_oracle.c -- FP32 CPU restatement (test infrastructure only) of
 * contact generation.  What each part restates (paths relative to the Talus
 * tree; glm is the reference's pinned GLM 1.0.1, commit 0af55cce, whose
 * quaternion product, quaternion-vector rotation, cross product and mix are
 * the textbook formulas written out below in glm's evaluation order):
 *   field_probe / probe_distance / field_material
 *                         src/talus/field/sampledsdf/sampledsdf.cc:18-23,
 *                         234-305, 387-402, 417-422, 432-477
 *   orc_field_lipschitz   sampledsdf.cc:61-93 (cellLipschitz), 115-214 and
 *                         584-606 (the global bound of a finished field),
 *                         424-430 (probeLipschitz)
 *   advance_shape ... field_point_of / probe_shape
 *                         src/talus/deformable/contact/contact.cc:10-49
 *   barycentric           contact.cc:87-156 (closest point on a triangle by
 *                         its Voronoi regions)
 *   closest_on            contact.cc:159-297 (seven starts, projected descent)
 *   swept_bounds ... advance_to_impact
 *                         contact.cc:299-536
 *   orc_find_contacts     contact.cc:599-880
 * Built with -ffp-contract=off like the reference (CMakeLists.txt:65), so
 * every float operation rounds as the reference's does; sinf and cosf are
 * the C library's, as glm::angleAxis calls them.  Pinned by
 * tests/test_contact_cpu.py: the reference's probe test of a sampled sphere
 * (field_meshsdf_tests.cc:391-415), exact planes, brute-force closest
 * points, the closed-form impact of a falling and a rising body, culling
 * that changes nothing, and the friction mix. */
#include "contact_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ADVANCE_ITERATIONS 4u
#define DESCENT_ITERATIONS 8u
#define BACKTRACKS 4u
#define IMPACT_RELATIVE_ERROR 0.05f
#define CULL_TOLERANCE 1.0e-5f
#define GRADIENT_FLOOR 1e-6f
#define DEFAULT_FRICTION 0.5f

typedef struct {
  float x, y, z;
} v3;
typedef struct {
  float w, x, y, z;
} q4;

static v3 V(float x, float y, float z) {
  v3 r = {x, y, z};
  return r;
}
static v3 load3(const float *p) { return V(p[0], p[1], p[2]); }
static void store3(v3 a, float *p) {
  p[0] = a.x;
  p[1] = a.y;
  p[2] = a.z;
}
static v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 scale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static v3 divide(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float length(v3 a) { return sqrtf(dot(a, a)); }
static v3 cross(v3 a, v3 b) { return V(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
static float mn(float a, float b) { return b < a ? b : a; }
static float mx(float a, float b) { return a < b ? b : a; }
static v3 vmin(v3 a, v3 b) { return V(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
static v3 vmax(v3 a, v3 b) { return V(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }
static float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }

/* Hamilton product */
static q4 qmul(q4 p, q4 q) {
  q4 r;
  r.w = p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z;
  r.x = p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y;
  r.y = p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z;
  r.z = p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x;
  return r;
}
static q4 qconj(q4 q) {
  q4 r = {q.w, -q.x, -q.y, -q.z};
  return r;
}
/* q v q*: v + 2 (w (u x v) + u x (u x v)), u the vector part */
static v3 rotate(q4 q, v3 v) {
  const v3 u = V(q.x, q.y, q.z);
  const v3 uv = cross(u, v);
  const v3 uuv = cross(u, uv);
  return add(v, scale(add(scale(uv, q.w), uuv), 2.0f));
}

/* ------------------------------------------------------------- the field */

static size_t at_index(const orc_sdf *f, uint32_t x, uint32_t y, uint32_t z) {
  return x + (size_t)f->res[0] * (y + (size_t)f->res[1] * z);
}

static v3 field_low(const orc_sdf *f) { return load3(f->origin); }
static v3 field_high(const orc_sdf *f) {
  return V(f->origin[0] + f->voxel[0] * (float)(f->res[0] - 1u), f->origin[1] + f->voxel[1] * (float)(f->res[1] - 1u),
           f->origin[2] + f->voxel[2] * (float)(f->res[2] - 1u));
}

typedef struct {
  uint32_t base[3];
  float t[3];
} cell_t;

static cell_t locate(const orc_sdf *f, v3 p) {
  cell_t c;
  const float lattice[3] = {(p.x - f->origin[0]) / f->voxel[0], (p.y - f->origin[1]) / f->voxel[1],
                            (p.z - f->origin[2]) / f->voxel[2]};
  for (int k = 0; k < 3; ++k) {
    uint32_t b = (uint32_t)floorf(lattice[k]);
    if (b > f->res[k] - 2u) b = f->res[k] - 2u;
    c.base[k] = b;
    c.t[k] = lattice[k] - (float)b;
  }
  return c;
}

static void corners(const orc_sdf *f, const cell_t *c, float *v) {
  for (uint32_t i = 0; i < 8u; ++i)
    v[i] = f->dist[at_index(f, c->base[0] + (i & 1u), c->base[1] + ((i >> 1) & 1u), c->base[2] + ((i >> 2) & 1u))];
}

static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

static float interpolate(const orc_sdf *f, const cell_t *c) {
  float v[8];
  corners(f, c, v);
  const float x00 = mixf(v[0], v[1], c->t[0]);
  const float x10 = mixf(v[2], v[3], c->t[0]);
  const float x01 = mixf(v[4], v[5], c->t[0]);
  const float x11 = mixf(v[6], v[7], c->t[0]);
  const float y0 = mixf(x00, x10, c->t[1]);
  const float y1 = mixf(x01, x11, c->t[1]);
  return mixf(y0, y1, c->t[2]);
}

static v3 slope(const orc_sdf *f, const cell_t *cell) {
  float c[8];
  corners(f, cell, c);
  const float u = cell->t[0], v = cell->t[1], w = cell->t[2];
  const v3 d = V((c[1] - c[0]) * (1.0f - v) * (1.0f - w) + (c[3] - c[2]) * v * (1.0f - w) +
                     (c[5] - c[4]) * (1.0f - v) * w + (c[7] - c[6]) * v * w,
                 (c[2] - c[0]) * (1.0f - u) * (1.0f - w) + (c[3] - c[1]) * u * (1.0f - w) +
                     (c[6] - c[4]) * (1.0f - u) * w + (c[7] - c[5]) * u * w,
                 (c[4] - c[0]) * (1.0f - u) * (1.0f - v) + (c[5] - c[1]) * u * (1.0f - v) +
                     (c[6] - c[2]) * (1.0f - u) * v + (c[7] - c[3]) * u * v);
  return V(d.x / f->voxel[0], d.y / f->voxel[1], d.z / f->voxel[2]);
}

typedef struct {
  v3 clamped, gap;
  float outside;
} grid_clamp_t;

static grid_clamp_t clamp_to_grid(const orc_sdf *f, v3 p) {
  grid_clamp_t g;
  const v3 lo = field_low(f), hi = field_high(f);
  g.clamped = vmin(vmax(p, lo), hi);
  g.gap = sub(p, g.clamped);
  g.outside = length(g.gap);
  return g;
}

typedef struct {
  float distance;
  v3 normal;
} probe_t;

static probe_t field_probe(const orc_sdf *f, v3 point) {
  const grid_clamp_t at = clamp_to_grid(f, point);
  const cell_t cell = locate(f, at.clamped);
  const float inner = interpolate(f, &cell);
  probe_t r;
  r.distance = at.outside <= 0.0f ? inner : inner + at.outside;
  const v3 derivative = slope(f, &cell);
  const float steepness = length(derivative);
  const int usable = steepness > GRADIENT_FLOOR;
  if (at.outside <= 0.0f) {
    r.normal = usable ? divide(derivative, steepness) : V(0.0f, 1.0f, 0.0f);
    return r;
  }
  const v3 blended = add(usable ? scale(divide(derivative, steepness), fabsf(inner)) : V(0.0f, 0.0f, 0.0f), at.gap);
  const float reach = length(blended);
  r.normal = reach > GRADIENT_FLOOR ? divide(blended, reach) : divide(at.gap, at.outside);
  return r;
}

static float probe_distance(const orc_sdf *f, v3 point) {
  const grid_clamp_t at = clamp_to_grid(f, point);
  const cell_t cell = locate(f, at.clamped);
  const float inner = interpolate(f, &cell);
  return at.outside <= 0.0f ? inner : inner + at.outside;
}

static uint32_t field_material(const orc_sdf *f, v3 point) {
  const grid_clamp_t at = clamp_to_grid(f, point);
  if (at.outside > 0.0f || !f->mat) return 0u;
  const cell_t c = locate(f, point);
  return f->mat[at_index(f, c.base[0] + (c.t[0] < 0.5f ? 0u : 1u), c.base[1] + (c.t[1] < 0.5f ? 0u : 1u),
                         c.base[2] + (c.t[2] < 0.5f ? 0u : 1u))];
}

float orc_field_lipschitz(const orc_sdf *f) {
  float worst = 0.0f;
  for (uint32_t z = 0; z + 1u < f->res[2]; ++z)
    for (uint32_t y = 0; y + 1u < f->res[1]; ++y)
      for (uint32_t x = 0; x + 1u < f->res[0]; ++x) {
        float c[8];
        for (uint32_t i = 0; i < 8u; ++i) c[i] = f->dist[at_index(f, x + (i & 1u), y + ((i >> 1) & 1u), z + ((i >> 2) & 1u))];
        v3 s = V(0.0f, 0.0f, 0.0f);
        for (uint32_t i = 0; i < 8u; ++i) {
          if (!(i & 1u)) s.x = mx(s.x, fabsf(c[i | 1u] - c[i]));
          if (!(i & 2u)) s.y = mx(s.y, fabsf(c[i | 2u] - c[i]));
          if (!(i & 4u)) s.z = mx(s.z, fabsf(c[i | 4u] - c[i]));
        }
        worst = mx(worst, length(V(s.x / f->voxel[0], s.y / f->voxel[1], s.z / f->voxel[2])));
      }
  return sqrtf(worst * worst + 1.0f);
}

void orc_field_probe(const orc_sdf *f, const float *point, float *distance, float *normal) {
  const probe_t p = field_probe(f, load3(point));
  *distance = p.distance;
  store3(p.normal, normal);
}

float orc_field_probe_distance(const orc_sdf *f, const float *point) { return probe_distance(f, load3(point)); }

uint32_t orc_field_material(const orc_sdf *f, const float *point) { return field_material(f, load3(point)); }

/* ------------------------------------------------------------- the shape */

typedef struct {
  v3 center;
  q4 rotation;
} shape_at_t;

static q4 shape_rotation(const orc_shape *s) {
  q4 q = {s->rotation[0], s->rotation[1], s->rotation[2], s->rotation[3]};
  return q;
}

static shape_at_t advance_shape(const orc_shape *s, float elapsed) {
  shape_at_t at;
  at.center = add(load3(s->center), scale(load3(s->linear), elapsed));
  const v3 w = load3(s->angular);
  const float spin = length(w);
  if (spin > 0.0f) {
    const v3 axis = divide(w, spin);
    const float half = spin * elapsed * 0.5f;
    const float sn = sinf(half);
    const q4 turn = {cosf(half), axis.x * sn, axis.y * sn, axis.z * sn};
    at.rotation = qmul(turn, shape_rotation(s));
  } else {
    at.rotation = shape_rotation(s);
  }
  return at;
}

void orc_advance_shape(const orc_shape *s, float elapsed, float *center, float *rotation) {
  const shape_at_t at = advance_shape(s, elapsed);
  store3(at.center, center);
  rotation[0] = at.rotation.w;
  rotation[1] = at.rotation.x;
  rotation[2] = at.rotation.y;
  rotation[3] = at.rotation.z;
}

static v3 field_point_of(const orc_shape *s, const shape_at_t *at, v3 world) {
  return add(rotate(qconj(at->rotation), sub(world, at->center)), load3(s->offset));
}

static probe_t probe_shape(const orc_shape *s, const shape_at_t *at, v3 world) {
  const probe_t probed = field_probe(s->field, field_point_of(s, at, world));
  const v3 outward = rotate(at->rotation, probed.normal);
  const float l = length(outward);
  probe_t r;
  r.distance = probed.distance;
  r.normal = l > 0.0f ? divide(outward, l) : V(0.0f, 1.0f, 0.0f);
  return r;
}

/* ---------------------------------------------------------- the triangle */

typedef struct {
  v3 c[3];
} tri_t;

static v3 barycentric(const tri_t *t, v3 p) {
  const v3 a = t->c[0], b = t->c[1], c = t->c[2];
  const v3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return V(1.0f, 0.0f, 0.0f);
  const v3 bp = sub(p, b);
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return V(0.0f, 1.0f, 0.0f);
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = d1 / (d1 - d3);
    return V(1.0f - v, v, 0.0f);
  }
  const v3 cp = sub(p, c);
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return V(0.0f, 0.0f, 1.0f);
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
    return V(1.0f - w, 0.0f, w);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return V(0.0f, 1.0f - w, w);
  }
  const float denominator = 1.0f / (va + vb + vc);
  const float v = vb * denominator, w = vc * denominator;
  return V(1.0f - v - w, v, w);
}

static v3 point_at(const tri_t *t, v3 w) {
  return add(add(scale(t->c[0], w.x), scale(t->c[1], w.y)), scale(t->c[2], w.z));
}

static float longest_edge(const tri_t *t) {
  return mx(length(sub(t->c[1], t->c[0])), mx(length(sub(t->c[2], t->c[1])), length(sub(t->c[0], t->c[2]))));
}

typedef struct {
  v3 weights, point;
  probe_t probe;
} closest_t;

static closest_t closest_on(const tri_t *t, const orc_shape *s, const shape_at_t *at) {
  const float third = 1.0f / 3.0f;
  const v3 starts[7] = {V(1.0f, 0.0f, 0.0f), V(0.0f, 1.0f, 0.0f), V(0.0f, 0.0f, 1.0f), V(0.5f, 0.5f, 0.0f),
                        V(0.0f, 0.5f, 0.5f), V(0.5f, 0.0f, 0.5f), V(third, third, third)};
  closest_t best;
  memset(&best, 0, sizeof best);
  best.probe.distance = FLT_MAX;
  best.probe.normal = V(0.0f, 1.0f, 0.0f);
  for (int i = 0; i < 7; ++i) {
    closest_t c;
    c.weights = starts[i];
    c.point = point_at(t, c.weights);
    c.probe = probe_shape(s, at, c.point);
    if (c.probe.distance < best.probe.distance) best = c;
  }
  if (!(best.probe.distance < FLT_MAX)) return best;
  const v3 across = cross(sub(t->c[1], t->c[0]), sub(t->c[2], t->c[0]));
  const float area = length(across);
  if (area <= 0.0f) return best;
  const v3 facing = divide(across, area);
  const float edge = longest_edge(t);
  for (uint32_t it = 0; it != DESCENT_ITERATIONS; ++it) {
    const v3 g = best.probe.normal;
    const v3 slide = neg(sub(g, scale(facing, dot(facing, g))));
    const float l = length(slide);
    if (l <= 1.0e-4f) break;
    const v3 direction = divide(slide, l);
    float stride = clampf(fabsf(best.probe.distance), 0.05f * edge, 0.5f * edge);
    int improved = 0;
    for (uint32_t a = 0; a != BACKTRACKS; ++a) {
      closest_t c;
      c.weights = barycentric(t, add(best.point, scale(direction, stride)));
      c.point = point_at(t, c.weights);
      c.probe = probe_shape(s, at, c.point);
      if (c.probe.distance < best.probe.distance) {
        best = c;
        improved = 1;
        break;
      }
      stride *= 0.5f;
    }
    if (!improved) break;
  }
  return best;
}

static tri_t tri_of(const float *p) {
  tri_t t;
  for (int c = 0; c < 3; ++c) t.c[c] = load3(p + 3 * c);
  return t;
}

void orc_barycentric(const float *triangle, const float *point, float *weights) {
  const tri_t t = tri_of(triangle);
  store3(barycentric(&t, load3(point)), weights);
}

void orc_closest_on_triangle(const orc_shape *s, float elapsed, const float *triangle, float *weights, float *point,
                             float *distance, float *normal) {
  const tri_t t = tri_of(triangle);
  const shape_at_t at = advance_shape(s, elapsed);
  const closest_t c = closest_on(&t, s, &at);
  store3(c.weights, weights);
  store3(c.point, point);
  *distance = c.probe.distance;
  store3(c.probe.normal, normal);
}

/* ------------------------------------------------------------ the bounds */

typedef struct {
  v3 lo, hi;
} bounds_t;

static bounds_t empty_bounds(void) {
  bounds_t b = {V(FLT_MAX, FLT_MAX, FLT_MAX), V(-FLT_MAX, -FLT_MAX, -FLT_MAX)};
  return b;
}
static void include(bounds_t *b, v3 p) {
  b->lo = vmin(b->lo, p);
  b->hi = vmax(b->hi, p);
}
static int overlaps(const bounds_t *a, const bounds_t *b, float margin) {
  return a->lo.x - margin <= b->hi.x && a->lo.y - margin <= b->hi.y && a->lo.z - margin <= b->hi.z &&
         b->lo.x - margin <= a->hi.x && b->lo.y - margin <= a->hi.y && b->lo.z - margin <= a->hi.z;
}

static bounds_t swept_bounds(const orc_shape *s, float dt) {
  const v3 low = sub(field_low(s->field), load3(s->offset));
  const v3 high = sub(field_high(s->field), load3(s->offset));
  bounds_t b = empty_bounds();
  const float times[2] = {0.0f, dt};
  for (int i = 0; i < 2; ++i) {
    const shape_at_t at = advance_shape(s, times[i]);
    for (uint32_t c = 0; c != 8u; ++c) {
      const v3 local = V((c & 1u) ? high.x : low.x, (c & 2u) ? high.y : low.y, (c & 4u) ? high.z : low.z);
      include(&b, add(at.center, rotate(at.rotation, local)));
    }
  }
  return b;
}

static float shape_radius(const orc_shape *s) {
  const v3 low = sub(field_low(s->field), load3(s->offset));
  const v3 high = sub(field_high(s->field), load3(s->offset));
  return length(vmax(V(fabsf(low.x), fabsf(low.y), fabsf(low.z)), V(fabsf(high.x), fabsf(high.y), fabsf(high.z))));
}

static float impact_bound(const orc_shape *s, float speed, float radius) {
  return speed + length(load3(s->linear)) + length(load3(s->angular)) * radius;
}

static float distance_below(const orc_shape *s, const shape_at_t *origin, v3 centre, float spread, float lipschitz) {
  const v3 local = field_point_of(s, origin, centre);
  const float sampled = probe_distance(s->field, local);
  const float magnitude = 1.0f + length(centre) + length(origin->center) + length(local) +
                          length(field_low(s->field)) + fabsf(sampled) + spread;
  return sampled - lipschitz * spread - CULL_TOLERANCE * (1.0f + lipschitz) * magnitude;
}

static int out_of_reach(const orc_shape *s, float lower, float speed, float dt, float bias, float threshold,
                        float radius) {
  if (!(lower > threshold)) return 0;
  const float bound = impact_bound(s, speed, radius);
  return bound <= 0.0f || (lower - bias) / bound > dt;
}

typedef struct {
  tri_t start;
  v3 velocity[3];
} moving_t;

static tri_t moving_at(const moving_t *m, float elapsed) {
  tri_t t = m->start;
  for (int c = 0; c < 3; ++c) t.c[c] = add(t.c[c], scale(m->velocity[c], elapsed));
  return t;
}

static float moving_speed(const moving_t *m) {
  float fastest = 0.0f;
  for (int c = 0; c < 3; ++c) fastest = mx(fastest, length(m->velocity[c]));
  return fastest;
}

static int far_from_shape(const orc_shape *s, const shape_at_t *origin, const moving_t *m, float dt, float bias,
                          float threshold, float radius, float lipschitz) {
  const tri_t *t = &m->start;
  const v3 centre = divide(add(add(t->c[0], t->c[1]), t->c[2]), 3.0f);
  float spread = 0.0f;
  for (int c = 0; c < 3; ++c) spread = mx(spread, length(sub(centre, t->c[c])));
  return out_of_reach(s, distance_below(s, origin, centre, spread, lipschitz), moving_speed(m), dt, bias, threshold,
                      radius);
}

typedef struct {
  bounds_t reach;
  v3 centre;
  float spread, speed, thickness;
} cluster_t;

static int cluster_far(const orc_shape *s, const shape_at_t *origin, const bounds_t *swept, const cluster_t *c,
                       float dt, float bias, float threshold, float radius, float lipschitz) {
  const float sb = bias + c->thickness, st = threshold + c->thickness;
  if (!overlaps(&c->reach, swept, st)) return 1;
  return out_of_reach(s, distance_below(s, origin, c->centre, c->spread, lipschitz), c->speed, dt, sb, st, radius);
}

static int advance_to_impact(const orc_shape *s, const moving_t *m, float dt, float bias, float radius,
                             closest_t *closest, float *impact_time) {
  const float bound = impact_bound(s, moving_speed(m), radius);
  if (bound <= 0.0f) return 0;
  float elapsed = (closest->probe.distance - bias) / bound;
  if (elapsed > dt) return 0;
  const float reach = bias * (1.0f + IMPACT_RELATIVE_ERROR);
  for (uint32_t it = 0; it != ADVANCE_ITERATIONS; ++it) {
    const shape_at_t at = advance_shape(s, elapsed);
    const tri_t t = moving_at(m, elapsed);
    *closest = closest_on(&t, s, &at);
    *impact_time = elapsed;
    if (closest->probe.distance <= reach) return 1;
    elapsed += (closest->probe.distance - bias) / bound;
    if (elapsed > dt) return 0;
  }
  return 1;
}

/* --------------------------------------------------------- findContacts */

static float friction_of(const orc_contact_mesh *m, uint32_t lane) {
  return lane < m->lanes ? m->friction[lane] : DEFAULT_FRICTION;
}

uint32_t orc_find_contacts(const orc_contact_mesh *m, const float *position, const float *velocity,
                           const float *gravity, const orc_shape *shapes, uint32_t nshapes, float dt, float bias,
                           float threshold, int cull, orc_surface_contact *out, uint32_t capacity,
                           orc_contact_stats *stats) {
  orc_contact_stats st;
  memset(&st, 0, sizeof st);
  const uint32_t faces = m->triangles;
  uint32_t count = 0;
  if (!nshapes || !faces) {
    if (stats) *stats = st;
    return 0;
  }
  const v3 g = load3(gravity);
  v3 *vel = (v3 *)malloc(sizeof(v3) * (m->nodes ? m->nodes : 1u));
  for (uint32_t n = 0; n < m->nodes; ++n)
    vel[n] = m->anchored && m->anchored[n] ? V(0.0f, 0.0f, 0.0f) : add(load3(velocity + 3 * (size_t)n), scale(g, dt));
  const float apart = 1.0e-2f * mx(mx(m->voxel[0], m->voxel[1]), m->voxel[2]);

  int clustered = cull && m->cluster_faces && m->cluster_starts && m->clusters >= 1u && m->cluster_starts[0] == 0u &&
                  m->cluster_starts[m->clusters] == faces;
  cluster_t *clusters = NULL;
  if (clustered) {
    clusters = (cluster_t *)calloc(m->clusters, sizeof(cluster_t));
    for (uint32_t k = 0; clustered && k != m->clusters; ++k) {
      cluster_t *h = &clusters[k];
      h->reach = empty_bounds();
      bounds_t start = empty_bounds();
      const uint32_t begin = m->cluster_starts[k], end = m->cluster_starts[k + 1u];
      if (end < begin) {
        clustered = 0;
        break;
      }
      for (uint32_t at = begin; at != end; ++at) {
        const uint32_t face = m->cluster_faces[at];
        if (face >= faces) {
          clustered = 0;
          break;
        }
        h->thickness = mx(h->thickness, m->thickness ? m->thickness[face] : 0.0f);
        for (int c = 0; c < 3; ++c) {
          const uint32_t node = m->tri[3 * (size_t)face + c];
          const v3 p = load3(position + 3 * (size_t)node);
          include(&start, p);
          include(&h->reach, p);
          include(&h->reach, add(p, scale(vel[node], dt)));
          h->speed = mx(h->speed, length(vel[node]));
        }
      }
      h->centre = scale(add(start.lo, start.hi), 0.5f);
      for (uint32_t at = begin; clustered && at != end; ++at) {
        const uint32_t face = m->cluster_faces[at];
        for (int c = 0; c < 3; ++c)
          h->spread = mx(h->spread, length(sub(h->centre, load3(position + 3 * (size_t)m->tri[3 * (size_t)face + c]))));
      }
    }
  }
  uint8_t *skipped = (uint8_t *)malloc(faces);
  v3 *points = (v3 *)malloc(sizeof(v3) * faces); /* mesh points of this shape's contacts */

  for (uint32_t index = 0; index != nshapes; ++index) {
    const orc_shape *s = &shapes[index];
    if (!s->field) continue;
    const bounds_t swept = swept_bounds(s, dt);
    const float radius = shape_radius(s);
    const shape_at_t origin = advance_shape(s, 0.0f);
    const float lipschitz = cull ? s->lipschitz : 0.0f;
    uint32_t held = 0;
    memset(skipped, 0, faces);
    if (clustered) {
      for (uint32_t k = 0; k != m->clusters; ++k) {
        if (!cluster_far(s, &origin, &swept, &clusters[k], dt, bias, threshold, radius, lipschitz)) continue;
        ++st.clusters_far;
        for (uint32_t at = m->cluster_starts[k]; at != m->cluster_starts[k + 1u]; ++at) skipped[m->cluster_faces[at]] = 1;
      }
    }
    for (uint32_t face = 0; face != faces; ++face) {
      if (clustered && skipped[face]) continue;
      const float thick = m->thickness ? m->thickness[face] : 0.0f;
      const float sb = bias + thick, sth = threshold + thick;
      moving_t mv;
      bounds_t reach = empty_bounds();
      for (int c = 0; c < 3; ++c) {
        const uint32_t node = m->tri[3 * (size_t)face + c];
        mv.start.c[c] = load3(position + 3 * (size_t)node);
        mv.velocity[c] = vel[node];
        include(&reach, mv.start.c[c]);
        include(&reach, add(mv.start.c[c], scale(mv.velocity[c], dt)));
      }
      if (!overlaps(&reach, &swept, sth)) {
        ++st.faces_apart;
        continue;
      }
      if (cull && far_from_shape(s, &origin, &mv, dt, sb, sth, radius, lipschitz)) {
        ++st.faces_far;
        continue;
      }
      ++st.faces_probed;
      closest_t closest = closest_on(&mv.start, s, &origin);
      float impact_time = 0.0f;
      if (closest.probe.distance > sth && !advance_to_impact(s, &mv, dt, sb, radius, &closest, &impact_time)) continue;
      const shape_at_t impact = advance_shape(s, impact_time);
      if (closest.probe.distance <= 0.0f) {
        const tri_t now = moving_at(&mv, impact_time);
        const v3 across = cross(sub(now.c[1], now.c[0]), sub(now.c[2], now.c[0]));
        const float area = length(across);
        if (area > 0.0f) closest.probe.normal = divide(neg(across), area);
      }
      const v3 on_shape = sub(closest.point, scale(closest.probe.normal, closest.probe.distance));
      const v3 local = rotate(qconj(impact.rotation), sub(on_shape, impact.center));
      const v3 mesh_point = point_at(&mv.start, closest.weights);
      const v3 shape_point = add(origin.center, rotate(origin.rotation, local));
      ++st.candidates;
      int duplicate = 0;
      for (uint32_t h = 0; h != held; ++h)
        if (length(sub(points[h], mesh_point)) <= apart) {
          duplicate = 1;
          break;
        }
      if (duplicate) {
        ++st.duplicates;
        continue;
      }
      points[held++] = mesh_point;
      if (count < capacity) {
        orc_surface_contact *c = &out[count];
        for (int k = 0; k < 3; ++k) c->nodes[k] = m->tri[3 * (size_t)face + k];
        store3(closest.weights, c->weights);
        c->partner = (int32_t)index;
        store3(closest.probe.normal, c->normal);
        store3(local, c->local_point);
        c->bias = sb;
        const uint32_t lane = m->lane ? m->lane[face] : 0u;
        const v3 fp = field_point_of(s, &impact, on_shape);
        c->friction = sqrtf(friction_of(m, lane) * friction_of(m, field_material(s->field, fp)));
        c->gap = dot(closest.probe.normal, sub(mesh_point, shape_point)) - sb;
        c->triangle = face;
      }
      ++count;
    }
  }
  free(points);
  free(skipped);
  free(clusters);
  free(vel);
  if (stats) *stats = st;
  return count;
}
