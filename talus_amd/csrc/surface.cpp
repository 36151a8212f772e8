// surface.cpp -- boundary surface of a tet mesh and its crack faces, on the
// host (collectSurface's tet faces, src/talus/deformable/solver/solver.cc:
// 220-364; markCracks, src/talus/dynamics/strainstep/tearing.cc:406-461,
// restated for tets).
//
// A face is on the surface when exactly one tet has it.  Faces are the
// corner triples {0,1,2}, {0,1,3}, {0,2,3}, {1,2,3} of every tet; a surface
// triangle keeps its tet's corner order, flipped when its normal points at
// the tet's fourth node, so every triangle faces outward.  Triangles come in
// ascending order of their sorted node triple, edges are the triangles'
// sorted node pairs without repeats, surface nodes the triangles' nodes in
// ascending order, and in clusters of 32 along the Morton order of their
// rest centres (clusterSurface, src/talus/deformable/contact/contact.cc:
// 540-597).  A crack face is a surface face whose three rest
// positions are those of another surface face: the two sides of a split made
// by a tear (split nodes keep their origin's rest position).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "host.h"

namespace tmi {

namespace {

constexpr int FACE[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

struct Face {
  std::array<uint32_t, 3> key;  // sorted node triple
  uint32_t tet;
  uint8_t which;
};

}  // namespace

void mesh_surface(uint32_t nodes, uint32_t tets, const uint32_t *tet, const uint32_t *lane,
                  const double *rest, SurfaceMesh &out) {
  out = SurfaceMesh{};
  out.crack.assign(tets, 0);
  std::vector<Face> faces;
  faces.reserve(4 * size_t(tets));
  for (uint32_t e = 0; e < tets; ++e)
    for (uint8_t f = 0; f < 4; ++f) {
      Face face{{tet[4 * size_t(e) + FACE[f][0]], tet[4 * size_t(e) + FACE[f][1]], tet[4 * size_t(e) + FACE[f][2]]}, e, f};
      std::sort(face.key.begin(), face.key.end());
      faces.push_back(face);
    }
  std::sort(faces.begin(), faces.end(), [](const Face &a, const Face &b) {
    return a.key != b.key ? a.key < b.key : (a.tet != b.tet ? a.tet < b.tet : a.which < b.which);
  });
  std::vector<const Face *> boundary;
  for (size_t at = 0; at < faces.size();) {
    size_t next = at + 1;
    while (next < faces.size() && faces[next].key == faces[at].key) ++next;
    if (next - at == 1) boundary.push_back(&faces[at]);
    at = next;
  }
  std::vector<uint8_t> on(nodes, 0);
  for (const Face *face : boundary) {
    const uint32_t *t = &tet[4 * size_t(face->tet)];
    uint32_t a = t[FACE[face->which][0]], b = t[FACE[face->which][1]], c = t[FACE[face->which][2]];
    const uint32_t d = t[6 - FACE[face->which][0] - FACE[face->which][1] - FACE[face->which][2]];
    double u[3], w[3], q[3];
    for (int k = 0; k < 3; ++k) {
      u[k] = rest[3 * size_t(b) + k] - rest[3 * size_t(a) + k];
      w[k] = rest[3 * size_t(c) + k] - rest[3 * size_t(a) + k];
      q[k] = rest[3 * size_t(d) + k] - rest[3 * size_t(a) + k];
    }
    const double n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    if (n[0] * q[0] + n[1] * q[1] + n[2] * q[2] > 0.0) std::swap(b, c);  // normal toward the tet: flip
    out.triangles.insert(out.triangles.end(), {a, b, c});
    out.triangle_tet.push_back(face->tet);
    out.triangle_lane.push_back(lane ? lane[face->tet] : 0u);
    const std::array<uint32_t, 3> &s = face->key;
    out.edges.insert(out.edges.end(), {s[0], s[1], s[0], s[2], s[1], s[2]});
    on[a] = on[b] = on[c] = 1;
  }
  // edges: sorted pairs without repeats
  std::vector<std::array<uint32_t, 2>> pairs(out.edges.size() / 2);
  for (size_t i = 0; i < pairs.size(); ++i) pairs[i] = {out.edges[2 * i], out.edges[2 * i + 1]};
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  out.edges.clear();
  for (const auto &p : pairs) out.edges.insert(out.edges.end(), {p[0], p[1]});
  for (uint32_t v = 0; v < nodes; ++v)
    if (on[v]) out.nodes.push_back(v);
  // crack faces: surface faces whose rest-position triples coincide
  struct Spot {
    std::array<double, 9> where;  // the three rest positions, sorted
    uint32_t tri;
  };
  std::vector<Spot> spots(boundary.size());
  for (size_t i = 0; i < boundary.size(); ++i) {
    std::array<std::array<double, 3>, 3> p;
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) p[k][r] = rest[3 * size_t(boundary[i]->key[k]) + r];
    std::sort(p.begin(), p.end());
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 3; ++r) spots[i].where[3 * k + r] = p[k][r];
    spots[i].tri = uint32_t(i);
  }
  std::sort(spots.begin(), spots.end(), [](const Spot &a, const Spot &b) {
    return a.where != b.where ? a.where < b.where : a.tri < b.tri;
  });
  for (size_t at = 0; at < spots.size();) {
    size_t next = at + 1;
    while (next < spots.size() && spots[next].where == spots[at].where) ++next;
    if (next - at > 1)
      for (size_t i = at; i < next; ++i) {
        const Face *face = boundary[spots[i].tri];
        out.crack[face->tet] = uint8_t(out.crack[face->tet] | (1u << face->which));
      }
    at = next;
  }
  cluster_surface(out.triangles, rest, out.cluster_faces, out.cluster_starts);
}

// 10 bits of v spread to every third bit (bit i to bit 3i)
static uint32_t spread_bits(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

void cluster_surface(const std::vector<uint32_t> &triangles, const double *rest, std::vector<uint32_t> &faces,
                     std::vector<uint32_t> &starts) {
  constexpr uint32_t SURFACE_CLUSTER = 32;  // contact.h:81
  const size_t T = triangles.size() / 3;
  faces.clear();
  starts.clear();
  if (T == 0) return;
  // the rest positions are the mesh's float samples: centres in float
  std::vector<std::array<float, 3>> centre(T);
  float low[3], high[3];
  for (int k = 0; k < 3; ++k) {
    low[k] = std::numeric_limits<float>::max();
    high[k] = std::numeric_limits<float>::lowest();
  }
  for (size_t f = 0; f < T; ++f)
    for (int k = 0; k < 3; ++k) {
      const float sum = float(rest[3 * size_t(triangles[3 * f]) + k]) + float(rest[3 * size_t(triangles[3 * f + 1]) + k]) +
                        float(rest[3 * size_t(triangles[3 * f + 2]) + k]);
      centre[f][k] = sum / 3.0f;
      low[k] = std::min(low[k], centre[f][k]);
      high[k] = std::max(high[k], centre[f][k]);
    }
  float span[3];
  for (int k = 0; k < 3; ++k) span[k] = std::max(high[k] - low[k], std::numeric_limits<float>::min());
  std::vector<std::pair<uint32_t, uint32_t>> keyed(T);
  for (size_t f = 0; f < T; ++f) {
    uint32_t cell[3];
    for (int k = 0; k < 3; ++k) {
      const float t = std::clamp((centre[f][k] - low[k]) / span[k], 0.0f, 1.0f);
      cell[k] = uint32_t(t * 1023.0f);
    }
    keyed[f] = {spread_bits(cell[0]) | (spread_bits(cell[1]) << 1) | (spread_bits(cell[2]) << 2), uint32_t(f)};
  }
  std::sort(keyed.begin(), keyed.end());
  faces.reserve(T);
  for (const auto &k : keyed) faces.push_back(k.second);
  for (size_t s = 0; s < T; s += SURFACE_CLUSTER) starts.push_back(uint32_t(s));
  starts.push_back(uint32_t(T));
}

void hex_surface(uint32_t nodes, uint32_t hexes, const uint32_t *hex, const uint32_t *lane, const double *rest,
                 SurfaceMesh &out) {
  // quad corners in order and a corner of the opposite face
  constexpr int QUAD[6][5] = {{0, 2, 6, 4, 1}, {1, 3, 7, 5, 0}, {0, 1, 5, 4, 2},
                              {2, 3, 7, 6, 0}, {0, 1, 3, 2, 4}, {4, 5, 7, 6, 0}};
  struct Tri {
    std::array<uint32_t, 3> key, ordered;
    uint32_t opposite, cell;
  };
  out = SurfaceMesh{};
  std::vector<Tri> tris;
  tris.reserve(12 * size_t(hexes));
  for (uint32_t c = 0; c < hexes; ++c)
    for (const auto &q : QUAD) {
      const uint32_t *h = &hex[8 * size_t(c)];
      const uint32_t quad[4] = {h[q[0]], h[q[1]], h[q[2]], h[q[3]]};
      uint32_t low = 0;
      for (uint32_t k = 1; k < 4; ++k)
        if (quad[k] < quad[low]) low = k;
      for (uint32_t half = 0; half < 2; ++half) {
        Tri t;
        t.ordered = {quad[low], quad[(low + 1 + half) % 4], quad[(low + 2 + half) % 4]};
        t.key = t.ordered;
        std::sort(t.key.begin(), t.key.end());
        t.opposite = h[q[4]];
        t.cell = c;
        tris.push_back(t);
      }
    }
  std::sort(tris.begin(), tris.end(), [](const Tri &a, const Tri &b) {
    return a.key != b.key ? a.key < b.key : (a.cell < b.cell);
  });
  auto at_rest = [&](uint32_t n, int k) { return float(rest[3 * size_t(n) + k]); };
  for (size_t at = 0; at < tris.size();) {
    size_t next = at + 1;
    while (next < tris.size() && tris[next].key == tris[at].key) ++next;
    if (next - at == 1) {
      const Tri &t = tris[at];
      const uint32_t a = t.ordered[0], b = t.ordered[1], c = t.ordered[2];
      float u[3], w[3], o[3];
      for (int k = 0; k < 3; ++k) {
        u[k] = at_rest(b, k) - at_rest(a, k);
        w[k] = at_rest(c, k) - at_rest(a, k);
        o[k] = at_rest(t.opposite, k) - at_rest(a, k);
      }
      const float n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      const bool inward = n[0] * o[0] + n[1] * o[1] + n[2] * o[2] > 0.0f;
      if (inward)
        out.triangles.insert(out.triangles.end(), {a, c, b});
      else
        out.triangles.insert(out.triangles.end(), {a, b, c});
      out.triangle_tet.push_back(t.cell);
      out.triangle_lane.push_back(lane ? lane[t.cell] : 0u);
      const auto &s = t.key;
      out.edges.insert(out.edges.end(), {s[0], s[1], s[0], s[2], s[1], s[2]});
    }
    at = next;
  }
  std::vector<std::array<uint32_t, 2>> pairs(out.edges.size() / 2);
  for (size_t i = 0; i < pairs.size(); ++i) pairs[i] = {out.edges[2 * i], out.edges[2 * i + 1]};
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  out.edges.clear();
  for (const auto &p : pairs) out.edges.insert(out.edges.end(), {p[0], p[1]});
  (void)nodes;
  cluster_surface(out.triangles, rest, out.cluster_faces, out.cluster_starts);
}

float sdf_distance(const talus_mi_sdf &f, const float *p) {
  float gap2 = 0.0f;
  float clamped[3];
  for (int k = 0; k < 3; ++k) {
    const float hi = f.origin[k] + f.voxel_size[k] * float(f.resolution[k] - 1u);
    clamped[k] = std::clamp(p[k], f.origin[k], hi);
    const float g = p[k] - clamped[k];
    gap2 += g * g;
  }
  const float outside = std::sqrt(gap2);
  if (outside > 0.0f) return outside;
  uint32_t base[3];
  float t[3];
  for (int k = 0; k < 3; ++k) {
    const float lattice = (p[k] - f.origin[k]) / f.voxel_size[k];
    base[k] = std::min(uint32_t(std::floor(lattice)), f.resolution[k] - 2u);
    t[k] = lattice - float(base[k]);
  }
  float c[8];
  for (uint32_t i = 0; i < 8; ++i) {
    const size_t x = base[0] + (i & 1u), y = base[1] + ((i >> 1) & 1u), z = base[2] + ((i >> 2) & 1u);
    c[i] = f.distances[x + size_t(f.resolution[0]) * (y + size_t(f.resolution[1]) * z)];
  }
  auto mix = [](float a, float b, float s) { return a * (1.0f - s) + b * s; };
  const float x00 = mix(c[0], c[1], t[0]), x10 = mix(c[2], c[3], t[0]);
  const float x01 = mix(c[4], c[5], t[0]), x11 = mix(c[6], c[7], t[0]);
  return mix(mix(x00, x10, t[1]), mix(x01, x11, t[1]), t[2]);
}

void surface_thickness(const talus_mi_sdf &field, const double *rest, uint32_t triangles,
                       const uint32_t *triangle_nodes, float *thickness) {
  for (uint32_t f = 0; f < triangles; ++f) {
    float sum = 0.0f;
    for (int c = 0; c < 3; ++c) {
      const double *r = &rest[3 * size_t(triangle_nodes[3 * size_t(f) + c])];
      const float p[3] = {float(r[0]), float(r[1]), float(r[2])};
      sum -= sdf_distance(field, p);
    }
    thickness[f] = sum / 3.0f;
  }
}

}  // namespace tmi
