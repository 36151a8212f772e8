// tear.cpp -- host topology update after tear events (tet version of
// src/talus/dynamics/strainstep/tearing.cc:180-513).
//
// For every torn tet, in ascending order (tearCells, tearing.cc:485-513):
//   crack normal n = top eigenvector of F^T F of the tet at the current
//   positions (stretchOf, 248-290; the sign is made canonical: largest
//   magnitude component positive, lowest axis on ties);
//   candidate nodes = the tet's corners sorted by |(X - centre) . n|, ties by
//   node id (crackThroughCell, 292-357); the first unpinned node whose fan
//   (tets containing it, ascending) has tets on both sides of the crack plane
//   (side = sum over the tet's corners of (X_c - X_node) . n > 0), and whose
//   face-connected groups per side each reach >= MIN_PIECE_CELLS = 2 tets
//   without crossing a face of the node (reachable, 131-165; with a limit of
//   2 that holds when a group has two tets, or when its one tet has a
//   face-neighbour through a face without the node), is split:
//   groups 1.. get a new node each (appended ids), group 0 keeps the node
//   (split/splitNode, 180-246);
// then nodes near the split ones whose fan falls apart are split the same
// way (splitPinched, 359-388) and masses are recomputed from the tets
// (recomputeMasses, 390-404).  New nodes copy rest position, position,
// velocity and acceleration of the node they were split from, are not
// anchored and start with zero mass before the recomputation.
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>
#include <set>
#include <vector>

#include "host.h"
#include "tmath.h"

namespace tmi {

namespace {

constexpr size_t MIN_PIECE_CELLS = 2;

using Face = std::array<uint32_t, 3>;

Face face_of(const uint32_t *t, int f) {
  // face f omits corner f
  Face out{};
  int w = 0;
  for (int c = 0; c < 4; ++c)
    if (c != f) out[w++] = t[c];
  std::sort(out.begin(), out.end());
  return out;
}

struct Topology {
  std::vector<uint32_t> &tet;  // 4 per tet
  uint32_t tets;
  std::vector<std::vector<uint32_t>> inc;  // node -> tets containing it, ascending

  Topology(std::vector<uint32_t> &t, uint32_t count, uint32_t nodes) : tet(t), tets(count), inc(nodes) {
    for (uint32_t e = 0; e < tets; ++e)
      for (int c = 0; c < 4; ++c) inc[tet[4 * size_t(e) + c]].push_back(e);
  }

  const uint32_t *nodes(uint32_t e) const { return &tet[4 * size_t(e)]; }

  const std::vector<uint32_t> &fan(uint32_t node) const { return inc[node]; }

  // moves tet e from `node` to the new node `copy`
  void relabel(uint32_t e, uint32_t node, uint32_t copy) {
    uint32_t *t = &tet[4 * size_t(e)];
    for (int c = 0; c < 4; ++c)
      if (t[c] == node) t[c] = copy;
    auto &from = inc[node];
    from.erase(std::lower_bound(from.begin(), from.end(), e));
    auto &to = inc[copy];
    to.insert(std::upper_bound(to.begin(), to.end(), e), e);
  }

  uint32_t shared(uint32_t a, uint32_t b) const {
    uint32_t n = 0;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) n += nodes(a)[i] == nodes(b)[j] ? 1u : 0u;
    return n;
  }

  // groupsOf (tearing.cc:92-129): same-side fan tets sharing a face join one
  // group; groups are numbered side 0 first, each side in fan order
  std::vector<uint32_t> groups(const std::vector<uint32_t> &fan, const std::vector<uint32_t> &sides) const {
    std::vector<uint32_t> parent(fan.size());
    std::iota(parent.begin(), parent.end(), 0u);
    auto root = [&](uint32_t x) {
      while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
      }
      return x;
    };
    for (uint32_t i = 0; i < fan.size(); ++i)
      for (uint32_t j = i + 1; j < fan.size(); ++j)
        if (sides[i] == sides[j] && shared(fan[i], fan[j]) >= 3u) {
          const uint32_t a = root(i), b = root(j);
          if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
    std::map<uint32_t, uint32_t> id;
    std::vector<uint32_t> out(fan.size(), 0u);
    for (uint32_t side = 0; side < 2; ++side)
      for (uint32_t i = 0; i < fan.size(); ++i) {
        if (sides[i] != side) continue;
        const uint32_t r = root(i);
        id.emplace(r, uint32_t(id.size()));
        out[i] = id.at(r);
      }
    return out;
  }

  // reachable(seeds, pivot, 2) >= 2
  bool whole(const std::vector<uint32_t> &seeds, uint32_t pivot) const {
    if (seeds.size() >= MIN_PIECE_CELLS) return true;
    const uint32_t e = seeds[0];
    for (int f = 0; f < 4; ++f) {
      const Face fc = face_of(nodes(e), f);
      if (fc[0] == pivot || fc[1] == pivot || fc[2] == pivot) continue;
      for (uint32_t o : inc[fc[0]]) {
        if (o == e) continue;
        const uint32_t *t = nodes(o);
        int hits = 0;
        for (int c = 0; c < 4; ++c) hits += (t[c] == fc[1] || t[c] == fc[2]) ? 1 : 0;
        if (hits == 2) return true;
      }
    }
    return false;
  }
};

void canonical(double *n) {
  int best = 0;
  for (int k = 1; k < 3; ++k)
    if (std::fabs(n[k]) > std::fabs(n[best])) best = k;
  if (n[best] < 0.0)
    for (int k = 0; k < 3; ++k) n[k] = -n[k];
}

}  // namespace

// Crack normal of tet e at positions x (FP64), canonical sign.
void tear_direction(const std::vector<double> &rest, const std::vector<double> &x,
                    const uint32_t *t, double *dir) {
  double Ds[9], ds[9], Dinv[9], F[9];
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) {
      Ds[r * 3 + k] = rest[3 * size_t(t[k + 1]) + r] - rest[3 * size_t(t[0]) + r];
      ds[r * 3 + k] = x[3 * size_t(t[k + 1]) + r] - x[3 * size_t(t[0]) + r];
    }
  inv3(Ds, Dinv);
  mul3(ds, Dinv, F);
  stretch_of(F, dir);
  canonical(dir);
}

uint32_t tear_topology(Body &b, const std::vector<uint32_t> &torn, std::vector<uint32_t> &origin) {
  origin.clear();
  if (torn.empty()) return 0;
  Topology topo(b.tet, b.tets, b.nodes);
  std::vector<std::pair<uint32_t, uint32_t>> copies;  // (copy, original)
  auto pinned = [&](uint32_t n) { return b.anchored[n] != 0; };
  auto split = [&](uint32_t node, const std::vector<uint32_t> &fan, const std::vector<uint32_t> &group) {
    const uint32_t count = group.empty() ? 0u : *std::max_element(group.begin(), group.end()) + 1u;
    if (count < 2 || pinned(node)) return false;
    for (uint32_t piece = 1; piece < count; ++piece) {
      const uint32_t copy = b.nodes++;
      for (int k = 0; k < 3; ++k) b.rest.push_back(b.rest[3 * size_t(node) + k]);
      for (int k = 0; k < 3; ++k) b.x.push_back(b.x[3 * size_t(node) + k]);
      for (int k = 0; k < 3; ++k) b.v.push_back(b.v[3 * size_t(node) + k]);
      for (int k = 0; k < 3; ++k) b.a.push_back(b.a[3 * size_t(node) + k]);
      b.mass.push_back(0.0);
      b.anchored.push_back(0);
      topo.inc.emplace_back();
      copies.emplace_back(copy, node);
      for (size_t i = 0; i < fan.size(); ++i)
        if (group[i] == piece) topo.relabel(fan[i], node, copy);
    }
    return true;
  };
  for (uint32_t cell : torn) {
    if (cell >= b.tets) continue;
    const uint32_t corners[4] = {b.tet[4 * size_t(cell)], b.tet[4 * size_t(cell) + 1],
                                 b.tet[4 * size_t(cell) + 2], b.tet[4 * size_t(cell) + 3]};
    double across[3];
    tear_direction(b.rest, b.x, corners, across);
    double centre[3] = {0, 0, 0};
    for (uint32_t n : corners)
      for (int k = 0; k < 3; ++k) centre[k] += b.rest[3 * size_t(n) + k] / 4.0;
    std::vector<std::pair<double, uint32_t>> order;
    for (uint32_t n : corners) {
      double d = 0.0;
      for (int k = 0; k < 3; ++k) d += (b.rest[3 * size_t(n) + k] - centre[k]) * across[k];
      order.emplace_back(std::fabs(d), n);
    }
    std::sort(order.begin(), order.end());
    for (const auto &[dist, node] : order) {
      (void)dist;
      if (pinned(node)) continue;
      const std::vector<uint32_t> fan = topo.fan(node);
      std::vector<uint32_t> sides(fan.size(), 0u);
      bool low = false, high = false;
      for (size_t i = 0; i < fan.size(); ++i) {
        double offset = 0.0;
        const uint32_t *t = topo.nodes(fan[i]);
        for (int c = 0; c < 4; ++c)
          for (int k = 0; k < 3; ++k) offset += (b.rest[3 * size_t(t[c]) + k] - b.rest[3 * size_t(node) + k]) * across[k];
        sides[i] = offset > 0.0 ? 1u : 0u;
        low = low || sides[i] == 0u;
        high = high || sides[i] == 1u;
      }
      if (!low || !high) continue;
      const std::vector<uint32_t> group = topo.groups(fan, sides);
      const uint32_t count = *std::max_element(group.begin(), group.end()) + 1u;
      bool whole = true;
      for (uint32_t piece = 0; piece < count && whole; ++piece) {
        std::vector<uint32_t> seeds;
        for (size_t i = 0; i < fan.size(); ++i)
          if (group[i] == piece) seeds.push_back(fan[i]);
        whole = topo.whole(seeds, node);
      }
      if (whole && split(node, fan, group)) break;
    }
  }
  if (copies.empty()) return 0;
  // splitPinched: nodes of tets touching a split node whose fan falls apart
  {
    std::set<uint32_t> near;
    for (const auto &[c, o] : copies) {
      near.insert(c);
      near.insert(o);
    }
    std::set<uint32_t> candidates;
    for (uint32_t n : near)
      for (uint32_t e : topo.fan(n)) candidates.insert(topo.nodes(e), topo.nodes(e) + 4);
    for (uint32_t node : candidates) {
      const std::vector<uint32_t> fan = topo.fan(node);
      if (fan.size() < 2) continue;
      (void)split(node, fan, topo.groups(fan, std::vector<uint32_t>(fan.size(), 0u)));
    }
  }
  for (const auto &pr : copies) origin.push_back(pr.second);
  return uint32_t(copies.size());
}

}  // namespace tmi
