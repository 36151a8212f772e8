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
// ascending order.  A crack face is a surface face whose three rest
// positions are those of another surface face: the two sides of a split made
// by a tear (split nodes keep their origin's rest position).
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
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
}

}  // namespace tmi
