/* field_oracle.c -- FP64 CPU restatement (test infrastructure only) of the
 * deformed-field re-bake.  What each part restates (paths relative to the
 * Talus tree):
 *   orc_sdf_probe        src/talus/field/sampledsdf/sampledsdf.cc:234-266
 *                        (locate, interpolate), 293-304 (clampToGrid),
 *                        432-440 (probe's distance)
 *   hex_shape / slope    src/talus/deformable/material/material.cc:310-347
 *   probe_build          contact.cc:1080-1150 (deformedProbe: cell keys,
 *                        boundary and crack faces)
 *   probe_place          contact.cc:1152-1285 (placeDeformedProbe)
 *   invert_hex           contact.cc:909-957 (invertHex)
 *   evaluate_hex         contact.cc:959-1027 (evaluateHex, distance only)
 *   screened / inside    contact.cc:1029-1078
 *   probe_deformed       contact.cc:1287-1425 (probeDeformed)
 *   orc_bake_deformed    src/talus/deformable/deformedfield/deformedfield.cc
 * Pinned by tests/test_field_cpu.py with the reference's deformed-field tests
 * (tests/deformedfield_tests.cc: at rest, moved and turned, a sunk top) and
 * closed forms of the trilinear probe and of a stretched lattice. */
#include "field_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DEFORMED_TOLERANCE 1.0e-10
#define DEFORMED_ITERATIONS 8
#define DEFORMED_PAD 1.5
#define DEFORMED_SCREEN 0.5
#define UNCHANGED 1.0e-7f
#define KEPT 16

static size_t sample_index(const orc_sdf *f, uint32_t x, uint32_t y, uint32_t z) {
  return x + (size_t)f->res[0] * (y + (size_t)f->res[1] * z);
}

double orc_sdf_probe(const orc_sdf *f, const double *p) {
  double clamped[3], gap2 = 0.0, t[3];
  uint32_t base[3];
  for (int k = 0; k < 3; ++k) {
    const double lo = (double)f->origin[k];
    const double hi = (double)(f->origin[k] + f->voxel[k] * (float)(f->res[k] - 1u));
    clamped[k] = p[k] < lo ? lo : (p[k] > hi ? hi : p[k]);
    gap2 += (p[k] - clamped[k]) * (p[k] - clamped[k]);
    const double lattice = (clamped[k] - (double)f->origin[k]) / (double)f->voxel[k];
    double fl = floor(lattice);
    uint32_t b = fl < 0.0 ? 0u : (uint32_t)fl;
    if (b > f->res[k] - 2u) b = f->res[k] - 2u;
    base[k] = b;
    t[k] = lattice - (double)b;
  }
  double c[8];
  for (uint32_t i = 0; i < 8; ++i)
    c[i] = (double)f->dist[sample_index(f, base[0] + (i & 1u), base[1] + ((i >> 1) & 1u), base[2] + ((i >> 2) & 1u))];
  /* glm::mix(a, b, t) = a * (1 - t) + b * t */
  const double x00 = c[0] * (1.0 - t[0]) + c[1] * t[0];
  const double x10 = c[2] * (1.0 - t[0]) + c[3] * t[0];
  const double x01 = c[4] * (1.0 - t[0]) + c[5] * t[0];
  const double x11 = c[6] * (1.0 - t[0]) + c[7] * t[0];
  const double y0 = x00 * (1.0 - t[1]) + x10 * t[1];
  const double y1 = x01 * (1.0 - t[1]) + x11 * t[1];
  const double inner = y0 * (1.0 - t[2]) + y1 * t[2];
  const double outside = sqrt(gap2);
  return outside <= 0.0 ? inner : inner + outside;
}

static void hex_shape(const double *local, double *shape) {
  for (uint32_t c = 0; c < 8; ++c) {
    double product = 1.0;
    for (int a = 0; a < 3; ++a) product *= ((c >> a) & 1u) ? local[a] : 1.0 - local[a];
    shape[c] = product;
  }
}

static void hex_slope(const double *local, double *slope) {
  for (uint32_t c = 0; c < 8; ++c) {
    double f[3], s[3];
    for (int a = 0; a < 3; ++a) {
      const int high = ((c >> a) & 1u) != 0u;
      f[a] = high ? local[a] : 1.0 - local[a];
      s[a] = high ? 1.0 : -1.0;
    }
    slope[3 * c + 0] = s[0] * f[1] * f[2];
    slope[3 * c + 1] = f[0] * s[1] * f[2];
    slope[3 * c + 2] = f[0] * f[1] * s[2];
  }
}

static double det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

static void inverse3(const double *a, double *o) {
  const double d = det3(a);
  o[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  o[1] = (a[2] * a[7] - a[1] * a[8]) / d;
  o[2] = (a[1] * a[5] - a[2] * a[4]) / d;
  o[3] = (a[5] * a[6] - a[3] * a[8]) / d;
  o[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  o[5] = (a[2] * a[3] - a[0] * a[5]) / d;
  o[6] = (a[3] * a[7] - a[4] * a[6]) / d;
  o[7] = (a[1] * a[6] - a[0] * a[7]) / d;
  o[8] = (a[0] * a[4] - a[1] * a[3]) / d;
}

typedef struct {
  const orc_hex_mesh *m;
  const orc_sdf *field;
  const double *pos;
  uint8_t *boundary, *cracked;
  double *low, *high, *centre, *inverse;
  double cell, pad_cells, pad;
  double minimum[3], maximum[3];
  uint32_t cells[3];
  uint32_t *first, *members;
} probe_t;

static uint64_t cell_key(const double *rest, const double *smallest, const double *voxel) {
  uint64_t key = 0;
  for (int k = 0; k < 3; ++k) {
    const double scaled = round((rest[k] - smallest[k]) / voxel[k]);
    const uint64_t packed = (uint64_t)((int64_t)scaled + (int64_t)(1ull << 20)) & 0x1fffffull;
    key |= packed << (21 * k);
  }
  return key;
}

static int cmp_u64(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static int has_key(const uint64_t *keys, size_t n, uint64_t key) {
  size_t lo = 0, hi = n;
  while (lo < hi) {
    const size_t mid = (lo + hi) / 2;
    if (keys[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && keys[lo] == key;
}

/* deformedProbe with infinite reach: boundary faces from the rest cells'
 * neighbours, crack faces added */
static void probe_build(probe_t *p, const orc_hex_mesh *m, const uint8_t *crack, const orc_sdf *field) {
  const size_t E = m->hexes;
  memset(p, 0, sizeof *p);
  p->m = m;
  p->field = field;
  float vmax = m->voxel[0];
  for (int k = 1; k < 3; ++k) vmax = m->voxel[k] > vmax ? m->voxel[k] : vmax;
  p->cell = (double)vmax;
  p->pad_cells = DEFORMED_PAD;
  p->pad = p->pad_cells * p->cell;
  double voxel[3], smallest[3] = {INFINITY, INFINITY, INFINITY};
  for (int k = 0; k < 3; ++k) voxel[k] = (double)m->voxel[k];
  for (size_t n = 0; n < m->nodes; ++n)
    for (int k = 0; k < 3; ++k) smallest[k] = m->rest[3 * n + k] < smallest[k] ? m->rest[3 * n + k] : smallest[k];
  uint64_t *keys = (uint64_t *)malloc((E + 1) * sizeof(uint64_t));
  for (size_t e = 0; e < E; ++e) keys[e] = cell_key(&m->rest[3 * (size_t)m->hex[8 * e]], smallest, voxel);
  qsort(keys, E, sizeof(uint64_t), cmp_u64);
  p->boundary = (uint8_t *)calloc(E + 1, 1);
  p->cracked = (uint8_t *)calloc(E + 1, 1);
  for (size_t e = 0; e < E; ++e) {
    const double *corner = &m->rest[3 * (size_t)m->hex[8 * e]];
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        double nb[3] = {corner[0], corner[1], corner[2]};
        nb[a] += side == 0 ? -voxel[a] : voxel[a];
        if (!has_key(keys, E, cell_key(nb, smallest, voxel))) p->boundary[e] |= (uint8_t)(1u << (2 * a + side));
      }
    if (crack) {
      p->cracked[e] = crack[e];
      p->boundary[e] |= crack[e];
    }
  }
  free(keys);
}

static void cell_of(const probe_t *p, const double *point, uint32_t *out) {
  for (int k = 0; k < 3; ++k) {
    double s = floor((point[k] - p->minimum[k]) / p->cell);
    s = s < 0.0 ? 0.0 : s;
    uint32_t c = (uint32_t)s;
    out[k] = c < p->cells[k] - 1u ? c : p->cells[k] - 1u;
  }
}

/* placeDeformedProbe (the unsized pad of an infinite reach) */
static void probe_place(probe_t *p, const double *position) {
  const orc_hex_mesh *m = p->m;
  const size_t E = m->hexes;
  p->pos = position;
  p->low = (double *)malloc((3 * E + 1) * sizeof(double));
  p->high = (double *)malloc((3 * E + 1) * sizeof(double));
  p->centre = (double *)malloc((3 * E + 1) * sizeof(double));
  p->inverse = (double *)calloc(9 * E + 1, sizeof(double));
  double middle[24];
  const double half[3] = {0.5, 0.5, 0.5};
  hex_slope(half, middle);
  for (size_t e = 0; e < E; ++e) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double c[3] = {0, 0, 0}, J[9] = {0};
    for (int k = 0; k < 8; ++k) {
      const double *at = &position[3 * (size_t)m->hex[8 * e + k]];
      for (int r = 0; r < 3; ++r) {
        lo[r] = at[r] < lo[r] ? at[r] : lo[r];
        hi[r] = at[r] > hi[r] ? at[r] : hi[r];
        c[r] += 0.125 * at[r];
        for (int s = 0; s < 3; ++s) J[3 * r + s] += at[r] * middle[3 * k + s];
      }
    }
    memcpy(&p->low[3 * e], lo, sizeof lo);
    memcpy(&p->high[3 * e], hi, sizeof hi);
    memcpy(&p->centre[3 * e], c, sizeof c);
    const double d = det3(J);
    if (d > 0.0 && isfinite(d)) inverse3(J, &p->inverse[9 * e]);
  }
  for (int k = 0; k < 3; ++k) {
    p->minimum[k] = INFINITY;
    p->maximum[k] = -INFINITY;
  }
  for (size_t e = 0; e < E; ++e) {
    const double reach = p->boundary[e] ? p->pad : 0.0;
    for (int k = 0; k < 3; ++k) {
      const double a = p->low[3 * e + k] - reach, b = p->high[3 * e + k] + reach;
      p->minimum[k] = a < p->minimum[k] ? a : p->minimum[k];
      p->maximum[k] = b > p->maximum[k] ? b : p->maximum[k];
    }
  }
  if (E == 0) {
    p->cells[0] = p->cells[1] = p->cells[2] = 0;
    p->first = (uint32_t *)calloc(1, sizeof(uint32_t));
    p->members = NULL;
    return;
  }
  for (int k = 0; k < 3; ++k) p->cells[k] = (uint32_t)floor((p->maximum[k] - p->minimum[k]) / p->cell) + 1u;
  const size_t total = (size_t)p->cells[0] * p->cells[1] * p->cells[2];
  p->first = (uint32_t *)calloc(total + 1, sizeof(uint32_t));
  for (int pass = 0; pass < 2; ++pass) {
    uint32_t *cursor = NULL;
    if (pass == 1) {
      for (size_t i = 0; i < total; ++i) p->first[i + 1] += p->first[i];
      p->members = (uint32_t *)malloc(((size_t)p->first[total] + 1) * sizeof(uint32_t));
      cursor = (uint32_t *)malloc((total + 1) * sizeof(uint32_t));
      memcpy(cursor, p->first, total * sizeof(uint32_t));
    }
    for (size_t e = 0; e < E; ++e) {
      const double reach = p->boundary[e] ? p->pad : 0.0;
      double a[3], b[3];
      uint32_t from[3], to[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = p->low[3 * e + k] - reach;
        b[k] = p->high[3 * e + k] + reach;
      }
      cell_of(p, a, from);
      cell_of(p, b, to);
      for (uint32_t z = from[2]; z <= to[2]; ++z)
        for (uint32_t y = from[1]; y <= to[1]; ++y)
          for (uint32_t x = from[0]; x <= to[0]; ++x) {
            const size_t index = x + p->cells[0] * ((size_t)y + (size_t)p->cells[1] * z);
            if (pass == 0)
              ++p->first[index + 1];
            else
              p->members[cursor[index]++] = (uint32_t)e;
          }
    }
    free(cursor);
  }
}

static void probe_free(probe_t *p) {
  free(p->boundary);
  free(p->cracked);
  free(p->low);
  free(p->high);
  free(p->centre);
  free(p->inverse);
  free(p->first);
  free(p->members);
}

typedef struct {
  double local[3];
  int converged, inverted;
} inversion_t;

static inversion_t invert_hex(const probe_t *p, uint32_t e, const double *point) {
  inversion_t f = {{0.5, 0.5, 0.5}, 0, 0};
  double determinant = 0.0;
  for (int it = 0; it < DEFORMED_ITERATIONS; ++it) {
    double shape[8], slope[24], mapped[3] = {0, 0, 0}, J[9] = {0}, Ji[9];
    hex_shape(f.local, shape);
    hex_slope(f.local, slope);
    for (int c = 0; c < 8; ++c) {
      const double *at = &p->pos[3 * (size_t)p->m->hex[8 * (size_t)e + c]];
      for (int r = 0; r < 3; ++r) {
        mapped[r] += shape[c] * at[r];
        for (int s = 0; s < 3; ++s) J[3 * r + s] += at[r] * slope[3 * c + s];
      }
    }
    determinant = det3(J);
    if (determinant == 0.0 || !isfinite(determinant)) break;
    inverse3(J, Ji);
    double step[3], most = 0.0;
    for (int r = 0; r < 3; ++r) {
      step[r] = Ji[3 * r] * (mapped[0] - point[0]) + Ji[3 * r + 1] * (mapped[1] - point[1]) +
                Ji[3 * r + 2] * (mapped[2] - point[2]);
      f.local[r] -= step[r];
      most = fabs(step[r]) > most ? fabs(step[r]) : most;
    }
    if (most < DEFORMED_TOLERANCE) {
      f.converged = 1;
      break;
    }
  }
  f.inverted = f.converged && !(determinant > 0.0);
  return f;
}

static double evaluate_hex(const probe_t *p, uint32_t e, const double *local) {
  double shape[8], slope[24], rest_point[3] = {0, 0, 0}, R[9] = {0};
  hex_shape(local, shape);
  hex_slope(local, slope);
  for (int c = 0; c < 8; ++c) {
    const double *rc = &p->m->rest[3 * (size_t)p->m->hex[8 * (size_t)e + c]];
    for (int r = 0; r < 3; ++r) {
      rest_point[r] += shape[c] * rc[r];
      for (int s = 0; s < 3; ++s) R[3 * r + s] += rc[r] * slope[3 * c + s];
    }
  }
  double distance = orc_sdf_probe(p->field, rest_point);
  const uint8_t cracks = p->cracked[e];
  for (int a = 0; a < 3 && cracks; ++a)
    for (int side = 0; side < 2; ++side) {
      if (!(cracks & (1u << (2 * a + side)))) continue;
      /* the length of the rest Jacobian's column a */
      const double extent = sqrt(R[a] * R[a] + R[3 + a] * R[3 + a] + R[6 + a] * R[6 + a]);
      const double depth = side == 0 ? local[a] : 1.0 - local[a];
      const double face = -depth * extent;
      if (face > distance) distance = face;
    }
  return distance;
}

static int screened(const probe_t *p, uint32_t e, const double *point, uint8_t faces) {
  const double *Mi = &p->inverse[9 * (size_t)e];
  int zero = 1;
  for (int k = 0; k < 9; ++k) zero = zero && Mi[k] == 0.0;
  if (zero) return 1;
  const double *c = &p->centre[3 * (size_t)e];
  const double d[3] = {point[0] - c[0], point[1] - c[1], point[2] - c[2]};
  for (int a = 0; a < 3; ++a) {
    const double value = Mi[3 * a] * d[0] + Mi[3 * a + 1] * d[1] + Mi[3 * a + 2] * d[2] + 0.5;
    if (value < -DEFORMED_SCREEN) {
      if (!(faces & (1u << (2 * a))) || value < -DEFORMED_SCREEN - p->pad_cells) return 0;
    } else if (value > 1.0 + DEFORMED_SCREEN) {
      if (!(faces & (1u << (2 * a + 1))) || value > 1.0 + DEFORMED_SCREEN + p->pad_cells) return 0;
    }
  }
  return 1;
}

static int inside(const double *point, const double *lo, const double *hi, double margin) {
  for (int k = 0; k < 3; ++k)
    if (!(point[k] >= lo[k] - margin && point[k] <= hi[k] + margin)) return 0;
  return 1;
}

/* probeDeformed: 1 with the distance when found */
static int probe_deformed(const probe_t *p, const double *point, double *distance, uint32_t *element,
                          double *local) {
  if (p->cells[0] == 0 || !inside(point, p->minimum, p->maximum, 0.0)) return 0;
  uint32_t cell[3];
  for (int k = 0; k < 3; ++k) {
    const uint32_t c = (uint32_t)floor((point[k] - p->minimum[k]) / p->cell);
    cell[k] = c < p->cells[k] - 1u ? c : p->cells[k] - 1u;
  }
  const size_t index = cell[0] + p->cells[0] * ((size_t)cell[1] + (size_t)p->cells[1] * cell[2]);
  const double margin = DEFORMED_TOLERANCE * p->cell;
  const uint32_t begin = p->first[index], end = p->first[index + 1];
  uint32_t kept_e[KEPT];
  inversion_t kept_i[KEPT];
  uint32_t kept = 0;
  for (uint32_t at = begin; at != end; ++at) {
    const uint32_t e = p->members[at];
    if (!inside(point, &p->low[3 * (size_t)e], &p->high[3 * (size_t)e], margin) || !screened(p, e, point, 0u))
      continue;
    const inversion_t f = invert_hex(p, e, point);
    if (kept != KEPT) {
      kept_e[kept] = e;
      kept_i[kept] = f;
      ++kept;
    }
    if (!f.converged || f.inverted) continue;
    const double *l = f.local;
    if (l[0] >= -DEFORMED_TOLERANCE && l[1] >= -DEFORMED_TOLERANCE && l[2] >= -DEFORMED_TOLERANCE &&
        l[0] <= 1.0 + DEFORMED_TOLERANCE && l[1] <= 1.0 + DEFORMED_TOLERANCE && l[2] <= 1.0 + DEFORMED_TOLERANCE) {
      *distance = evaluate_hex(p, e, l);
      if (element) *element = e;
      if (local) memcpy(local, l, 3 * sizeof(double));
      return 1;
    }
  }
  int accepted = 0;
  uint32_t chosen = 0;
  double chosen_local[3] = {0, 0, 0}, chosen_excess = INFINITY;
  for (uint32_t at = begin; at != end; ++at) {
    const uint32_t e = p->members[at];
    const uint8_t faces = p->boundary[e];
    if (faces == 0 || !inside(point, &p->low[3 * (size_t)e], &p->high[3 * (size_t)e], p->pad) ||
        !screened(p, e, point, faces))
      continue;
    inversion_t f;
    uint32_t hit = kept;
    for (uint32_t k = 0; k < kept; ++k)
      if (kept_e[k] == e) {
        hit = k;
        break;
      }
    f = hit != kept ? kept_i[hit] : invert_hex(p, e, point);
    if (!f.converged || f.inverted) continue;
    int across = 1;
    double excess = 0.0;
    for (int a = 0; a < 3 && across; ++a) {
      const double value = f.local[a];
      if (value < -DEFORMED_TOLERANCE) {
        across = (faces & (1u << (2 * a))) != 0;
        excess = -value > excess ? -value : excess;
      } else if (value > 1.0 + DEFORMED_TOLERANCE) {
        across = (faces & (1u << (2 * a + 1))) != 0;
        excess = value - 1.0 > excess ? value - 1.0 : excess;
      }
    }
    if (across && excess < chosen_excess) {
      accepted = 1;
      chosen = e;
      memcpy(chosen_local, f.local, sizeof chosen_local);
      chosen_excess = excess;
    }
  }
  if (!accepted) return 0;
  *distance = evaluate_hex(p, chosen, chosen_local);
  if (element) *element = chosen;
  if (local) memcpy(local, chosen_local, sizeof chosen_local);
  return 1;
}

int orc_probe_deformed(const orc_hex_mesh *mesh, const uint8_t *crack, const orc_sdf *rest,
                       const double *position, const double *point, double *distance, uint32_t *element,
                       double *local) {
  probe_t p;
  probe_build(&p, mesh, crack, rest);
  probe_place(&p, position);
  const int found = probe_deformed(&p, point, distance, element, local);
  probe_free(&p);
  return found;
}

int orc_bake_deformed(const orc_hex_mesh *mesh, const uint8_t *crack, const double *position, const double *model,
                      const orc_sdf *rest, orc_sdf *current, int carve, uint32_t *box_min, uint32_t *box_ext) {
  for (int k = 0; k < 3; ++k)
    if (current->res[k] != rest->res[k] || current->origin[k] != rest->origin[k] ||
        current->voxel[k] != rest->voxel[k])
      return 0;
  probe_t p;
  probe_build(&p, mesh, crack, rest);
  probe_place(&p, position);
  const size_t count = (size_t)rest->res[0] * rest->res[1] * rest->res[2];
  float *baked = (float *)malloc((count + 1) * sizeof(float));
  uint32_t lo[3] = {UINT32_MAX, UINT32_MAX, UINT32_MAX}, hi[3] = {0, 0, 0};
  int changed = 0;
  for (uint32_t z = 0; z < rest->res[2]; ++z)
    for (uint32_t y = 0; y < rest->res[1]; ++y)
      for (uint32_t x = 0; x < rest->res[0]; ++x) {
        const size_t index = sample_index(rest, x, y, z);
        const uint32_t xyz[3] = {x, y, z};
        double local[3], world[3], distance;
        for (int k = 0; k < 3; ++k) local[k] = (double)rest->origin[k] + (double)rest->voxel[k] * (double)xyz[k];
        for (int r = 0; r < 3; ++r)
          world[r] = model[4 * r] * local[0] + model[4 * r + 1] * local[1] + model[4 * r + 2] * local[2] + model[4 * r + 3];
        const float unplaced = carve ? (rest->dist[index] > (float)p.pad ? rest->dist[index] : (float)p.pad)
                                     : rest->dist[index];
        const float value = probe_deformed(&p, world, &distance, NULL, NULL) ? (float)distance : unplaced;
        if (fabsf(value - current->dist[index]) > UNCHANGED) {
          baked[index] = value;
          for (int k = 0; k < 3; ++k) {
            lo[k] = xyz[k] < lo[k] ? xyz[k] : lo[k];
            hi[k] = xyz[k] > hi[k] ? xyz[k] : hi[k];
          }
          changed = 1;
        } else {
          baked[index] = current->dist[index];
        }
      }
  memcpy(current->dist, baked, count * sizeof(float));
  for (int k = 0; k < 3; ++k) {
    box_min[k] = changed ? lo[k] : 0u;
    box_ext[k] = changed ? hi[k] - lo[k] + 1u : 0u;
  }
  free(baked);
  probe_free(&p);
  return 1;
}

float orc_sdf_distance(const orc_sdf *f, const float *p) {
  float gap2 = 0.0f;
  for (int k = 0; k < 3; ++k) {
    const float hi = f->origin[k] + f->voxel[k] * (float)(f->res[k] - 1u);
    const float c = p[k] < f->origin[k] ? f->origin[k] : (p[k] > hi ? hi : p[k]);
    const float g = p[k] - c;
    gap2 += g * g;
  }
  const float outside = sqrtf(gap2);
  if (outside > 0.0f) return outside;
  uint32_t base[3];
  float t[3];
  for (int k = 0; k < 3; ++k) {
    const float lattice = (p[k] - f->origin[k]) / f->voxel[k];
    uint32_t b = (uint32_t)floorf(lattice);
    b = b < f->res[k] - 2u ? b : f->res[k] - 2u;
    base[k] = b;
    t[k] = lattice - (float)b;
  }
  float c[8];
  for (uint32_t i = 0; i < 8; ++i)
    c[i] = f->dist[sample_index(f, base[0] + (i & 1u), base[1] + ((i >> 1) & 1u), base[2] + ((i >> 2) & 1u))];
  const float x00 = c[0] * (1.0f - t[0]) + c[1] * t[0];
  const float x10 = c[2] * (1.0f - t[0]) + c[3] * t[0];
  const float x01 = c[4] * (1.0f - t[0]) + c[5] * t[0];
  const float x11 = c[6] * (1.0f - t[0]) + c[7] * t[0];
  const float y0 = x00 * (1.0f - t[1]) + x10 * t[1];
  const float y1 = x01 * (1.0f - t[1]) + x11 * t[1];
  return y0 * (1.0f - t[2]) + y1 * t[2];
}

void orc_surface_thickness(const orc_sdf *f, const double *rest, uint32_t triangles, const uint32_t *tri,
                           float *thickness) {
  for (uint32_t t = 0; t < triangles; ++t) {
    float sum = 0.0f;
    for (int c = 0; c < 3; ++c) {
      const double *r = &rest[3 * (size_t)tri[3 * t + c]];
      const float p[3] = {(float)r[0], (float)r[1], (float)r[2]};
      sum -= orc_sdf_distance(f, p);
    }
    thickness[t] = sum / 3.0f;
  }
}

typedef struct {
  uint32_t key[3], ordered[3], opposite, cell;
} hex_tri_t;

static int cmp_hex_tri(const void *a, const void *b) {
  const hex_tri_t *x = (const hex_tri_t *)a, *y = (const hex_tri_t *)b;
  for (int k = 0; k < 3; ++k)
    if (x->key[k] != y->key[k]) return x->key[k] < y->key[k] ? -1 : 1;
  return x->cell < y->cell ? -1 : (x->cell > y->cell ? 1 : 0);
}

static int cmp_pair(const void *a, const void *b) {
  const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : (x[1] > y[1] ? 1 : 0);
}

static void sort3(uint32_t *v) {
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (v[j] < v[i]) {
        const uint32_t t = v[i];
        v[i] = v[j];
        v[j] = t;
      }
}

uint32_t orc_hex_surface(const orc_hex_mesh *m, uint32_t *tri, uint32_t *cell, uint32_t *lane, uint32_t *edges,
                         uint32_t *nedges) {
  static const int HEX_FACES[6][5] = {{0, 2, 6, 4, 1}, {1, 3, 7, 5, 0}, {0, 1, 5, 4, 2},
                                      {2, 3, 7, 6, 0}, {0, 1, 3, 2, 4}, {4, 5, 7, 6, 0}};
  const size_t H = m->hexes;
  hex_tri_t *t = (hex_tri_t *)malloc((12 * H + 1) * sizeof(hex_tri_t));
  size_t n = 0;
  for (size_t c = 0; c < H; ++c)
    for (int f = 0; f < 6; ++f) {
      const uint32_t *h = &m->hex[8 * c];
      const uint32_t quad[4] = {h[HEX_FACES[f][0]], h[HEX_FACES[f][1]], h[HEX_FACES[f][2]], h[HEX_FACES[f][3]]};
      int low = 0;
      for (int k = 1; k < 4; ++k)
        if (quad[k] < quad[low]) low = k;
      const int split[2][3] = {{low, (low + 1) % 4, (low + 2) % 4}, {low, (low + 2) % 4, (low + 3) % 4}};
      for (int s = 0; s < 2; ++s) {
        hex_tri_t *x = &t[n++];
        for (int k = 0; k < 3; ++k) x->ordered[k] = x->key[k] = quad[split[s][k]];
        sort3(x->key);
        x->opposite = h[HEX_FACES[f][4]];
        x->cell = (uint32_t)c;
      }
    }
  qsort(t, n, sizeof(hex_tri_t), cmp_hex_tri);
  uint32_t nt = 0, ne = 0;
  for (size_t at = 0; at < n;) {
    size_t next = at + 1;
    while (next < n && memcmp(t[next].key, t[at].key, sizeof t[at].key) == 0) ++next;
    if (next - at == 1) {
      const hex_tri_t *x = &t[at];
      const uint32_t a = x->ordered[0], b = x->ordered[1], c = x->ordered[2];
      float u[3], w[3], o[3];
      for (int k = 0; k < 3; ++k) {
        const float ra = (float)m->rest[3 * (size_t)a + k];
        u[k] = (float)m->rest[3 * (size_t)b + k] - ra;
        w[k] = (float)m->rest[3 * (size_t)c + k] - ra;
        o[k] = (float)m->rest[3 * (size_t)x->opposite + k] - ra;
      }
      const float nn[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      const int inward = nn[0] * o[0] + nn[1] * o[1] + nn[2] * o[2] > 0.0f;
      tri[3 * nt] = a;
      tri[3 * nt + 1] = inward ? c : b;
      tri[3 * nt + 2] = inward ? b : c;
      cell[nt] = x->cell;
      lane[nt] = m->lane ? m->lane[x->cell] : 0u;
      ++nt;
      const uint32_t *s = x->key;
      const uint32_t pairs[3][2] = {{s[0], s[1]}, {s[0], s[2]}, {s[1], s[2]}};
      for (int k = 0; k < 3; ++k) {
        edges[2 * ne] = pairs[k][0];
        edges[2 * ne + 1] = pairs[k][1];
        ++ne;
      }
    }
    at = next;
  }
  qsort(edges, ne, 2 * sizeof(uint32_t), cmp_pair);
  uint32_t kept = 0;
  for (uint32_t i = 0; i < ne; ++i)
    if (kept == 0 || edges[2 * i] != edges[2 * (kept - 1)] || edges[2 * i + 1] != edges[2 * (kept - 1) + 1]) {
      edges[2 * kept] = edges[2 * i];
      edges[2 * kept + 1] = edges[2 * i + 1];
      ++kept;
    }
  *nedges = kept;
  free(t);
  return nt;
}
