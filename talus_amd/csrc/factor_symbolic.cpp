// factor_symbolic.cpp -- host symbolic analysis of the block Cholesky
// preconditioner: nested dissection, supernodal fronts, element-pair
// assembly lists and the task lists of every factorization and solve launch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <stdexcept>

#include "factor.h"

namespace tmi {

namespace {

// node adjacency (nodes sharing a tet), CSR with sorted unique neighbours
void adjacency(const Body &b, std::vector<uint32_t> &start, std::vector<uint32_t> &adj) {
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  edges.reserve(size_t(b.tets) * 12);
  for (uint32_t e = 0; e < b.tets; ++e) {
    const uint32_t *t = &b.tet[4 * size_t(e)];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        if (i != j) edges.emplace_back(t[i], t[j]);
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  start.assign(b.nodes + 1, 0);
  adj.resize(edges.size());
  for (size_t k = 0; k < edges.size(); ++k) {
    ++start[edges[k].first + 1];
    adj[k] = edges[k].second;
  }
  for (uint32_t n = 0; n < b.nodes; ++n) start[n + 1] += start[n];
}

struct Dissector {
  const Body &b;
  const std::vector<uint32_t> &start, &adj;
  uint32_t leaf;
  std::vector<uint8_t> side;             // scratch marks
  std::vector<std::vector<uint32_t>> out;  // fronts' columns in postorder

  void emit(std::vector<uint32_t> &cols) {
    if (!cols.empty()) out.push_back(cols);
  }

  void run(std::vector<uint32_t> S) {
    if (S.size() <= leaf) {
      std::sort(S.begin(), S.end());
      emit(S);
      return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t n : S)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], b.rest[3 * size_t(n) + k]);
        hi[k] = std::max(hi[k], b.rest[3 * size_t(n) + k]);
      }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
      if (hi[k] - lo[k] > hi[axis] - lo[axis]) axis = k;
    std::stable_sort(S.begin(), S.end(), [&](uint32_t a, uint32_t c) {
      const double ca = b.rest[3 * size_t(a) + axis], cc = b.rest[3 * size_t(c) + axis];
      return ca < cc || (ca == cc && a < c);
    });
    const size_t n = S.size();
    size_t h = n / 2;
    const size_t w = std::max<size_t>(1, n / 8);
    const size_t lo_i = std::max<size_t>(1, h > w ? h - w : 1), hi_i = std::min(n - 1, h + w);
    double best = -1.0;
    for (size_t i = lo_i; i < hi_i; ++i) {
      const double gap = b.rest[3 * size_t(S[i]) + axis] - b.rest[3 * size_t(S[i - 1]) + axis];
      if (gap > best) {
        best = gap;
        h = i;
      }
    }
    for (size_t i = h; i < n; ++i) side[S[i]] = 1;
    std::vector<uint32_t> left, sep, right(S.begin() + h, S.end());
    for (size_t i = 0; i < h; ++i) {
      const uint32_t u = S[i];
      bool touches = false;
      for (uint32_t k = start[u]; k < start[u + 1] && !touches; ++k) touches = side[adj[k]] == 1;
      (touches ? sep : left).push_back(u);
    }
    for (size_t i = h; i < n; ++i) side[S[i]] = 0;
    S.clear();
    S.shrink_to_fit();
    if (!left.empty()) run(std::move(left));
    if (!right.empty()) run(std::move(right));
    std::sort(sep.begin(), sep.end());
    emit(sep);
  }
};

}  // namespace

Symbolic analyze(const Body &b, uint32_t leaf) {
  Symbolic s;
  std::vector<uint32_t> start, adj;
  adjacency(b, start, adj);
  std::vector<uint32_t> free_nodes;
  for (uint32_t n = 0; n < b.nodes; ++n)
    if (!b.anchored[n]) free_nodes.push_back(n);
  s.nfree = uint32_t(free_nodes.size());
  s.pos.assign(b.nodes, -1);
  if (free_nodes.empty()) return s;
  Dissector d{b, start, adj, std::max<uint32_t>(leaf, 1), std::vector<uint8_t>(b.nodes, 0), {}};
  d.run(free_nodes);
  // elimination order = emission order
  for (const auto &cols : d.out)
    for (uint32_t n : cols) {
      s.pos[n] = int32_t(s.perm.size());
      s.perm.push_back(n);
    }
  const uint32_t F = uint32_t(d.out.size());
  s.fronts.resize(F);
  s.front_of.resize(s.nfree);
  std::vector<uint32_t> first(F), last(F);
  {
    uint32_t e = 0;
    for (uint32_t f = 0; f < F; ++f) {
      first[f] = e;
      for (size_t i = 0; i < d.out[f].size(); ++i) s.front_of[e++] = f;
      last[f] = e - 1;
      s.fronts[f].ncol = uint32_t(d.out[f].size());
    }
  }
  std::vector<std::vector<uint32_t>> kids(F);
  std::vector<uint32_t> rowbuf;
  for (uint32_t f = 0; f < F; ++f) {
    rowbuf.clear();
    for (uint32_t e = first[f]; e <= last[f]; ++e) {
      const uint32_t u = s.perm[e];
      for (uint32_t k = start[u]; k < start[u + 1]; ++k) {
        const int32_t p = s.pos[adj[k]];
        if (p > int32_t(last[f])) rowbuf.push_back(uint32_t(p));
      }
    }
    for (uint32_t c : kids[f]) {
      const FrontSym &cf = s.fronts[c];
      for (uint32_t r = cf.ncol; r < cf.nrow; ++r) {
        const uint32_t p = s.rows[cf.row0 + r];
        if (p > last[f]) rowbuf.push_back(p);
      }
    }
    std::sort(rowbuf.begin(), rowbuf.end());
    rowbuf.erase(std::unique(rowbuf.begin(), rowbuf.end()), rowbuf.end());
    FrontSym &fs = s.fronts[f];
    fs.row0 = uint32_t(s.rows.size());
    for (uint32_t e = first[f]; e <= last[f]; ++e) s.rows.push_back(e);
    s.rows.insert(s.rows.end(), rowbuf.begin(), rowbuf.end());
    fs.nrow = fs.ncol + uint32_t(rowbuf.size());
    if (!rowbuf.empty()) {
      fs.parent = int32_t(s.front_of[rowbuf.front()]);
      kids[fs.parent].push_back(f);
    }
    uint32_t lv = 0;
    for (uint32_t c : kids[f]) lv = std::max(lv, s.fronts[c].level + 1);
    fs.level = lv;
    s.levels = std::max(s.levels, lv + 1);
    const double k = 3.0 * fs.ncol, m = 3.0 * fs.nrow;
    s.flops += k * k * k / 3.0 + k * k * (m - k) + k * (m - k) * (m - k) + k * k * k / 3.0;
    s.factor_entries += uint64_t(k * m - k * (k - 1) / 2);
  }
  s.child_start.assign(F + 1, 0);
  for (uint32_t f = 0; f < F; ++f) s.child_start[f + 1] = s.child_start[f] + uint32_t(kids[f].size());
  for (uint32_t f = 0; f < F; ++f) s.children.insert(s.children.end(), kids[f].begin(), kids[f].end());
  return s;
}

static inline uint32_t tiles(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

void build_tables(const std::vector<Symbolic> &sym, const std::vector<Body> &bodies,
                  const std::vector<uint32_t> &n0, const std::vector<uint32_t> &t0,
                  FactorTables &o) {
  o = FactorTables{};
  o.bodies = uint32_t(bodies.size());
  std::vector<uint32_t> fbase(bodies.size() + 1, 0);
  for (size_t b = 0; b < bodies.size(); ++b) {
    fbase[b + 1] = fbase[b] + uint32_t(sym[b].fronts.size());
    o.levels = std::max(o.levels, sym[b].levels);
    o.flops += sym[b].flops;
    o.factor_entries += sym[b].factor_entries;
  }
  o.fronts = fbase.back();
  const uint32_t F = o.fronts;
  o.k.resize(F);
  o.m.resize(F);
  o.body.resize(F);
  o.parent.resize(F);
  o.rows_off.resize(F);
  o.map_off.resize(F);
  o.level.resize(F);
  o.off.resize(F);
  o.minv_off.resize(F);
  o.vec_off.resize(F);
  o.yoff.resize(F);
  o.g_off.resize(F);
  o.child_start.assign(F + 1, 0);
  uint64_t pool = 0, minv = 0, vec = 0, yv = 0, gp = 0;
  for (size_t b = 0; b < bodies.size(); ++b) {
    const Symbolic &s = sym[b];
    for (uint32_t lf = 0; lf < s.fronts.size(); ++lf) {
      const uint32_t f = fbase[b] + lf;
      const FrontSym &fs = s.fronts[lf];
      o.k[f] = 3 * fs.ncol;
      o.m[f] = 3 * fs.nrow;
      o.body[f] = uint32_t(b);
      o.parent[f] = fs.parent < 0 ? ~0u : fbase[b] + uint32_t(fs.parent);
      o.level[f] = fs.level;
      o.off[f] = pool;
      pool += uint64_t(o.m[f]) * o.m[f];
      o.minv_off[f] = minv;
      minv += uint64_t(o.k[f]) * o.k[f];
      o.vec_off[f] = vec;
      vec += o.m[f];
      o.yoff[f] = yv;
      yv += o.k[f];
      o.g_off[f] = gp;
      gp += uint64_t(pad8(o.m[f])) * pad8(o.k[f]);  // both tiled copies fit, padding included
      o.rows_off[f] = uint32_t(o.rows.size());
      for (uint32_t r = 0; r < fs.nrow; ++r) o.rows.push_back(n0[b] + s.perm[s.rows[fs.row0 + r]]);
      const uint32_t cs = s.child_start[lf], ce = s.child_start[lf + 1];
      o.child_start[f + 1] = ce - cs;
      for (uint32_t c = cs; c < ce; ++c) o.children.push_back(fbase[b] + s.children[c]);
    }
    // update-row maps: row r (>= ncol) of front f -> node-row index in parent
    for (uint32_t lf = 0; lf < s.fronts.size(); ++lf) {
      const uint32_t f = fbase[b] + lf;
      const FrontSym &fs = s.fronts[lf];
      o.map_off[f] = uint32_t(o.map.size());
      if (fs.parent < 0) continue;
      const FrontSym &ps = s.fronts[fs.parent];
      const uint32_t *pr = &s.rows[ps.row0];
      for (uint32_t r = fs.ncol; r < fs.nrow; ++r) {
        const uint32_t e = s.rows[fs.row0 + r];
        const uint32_t *it = std::lower_bound(pr, pr + ps.nrow, e);
        if (it == pr + ps.nrow || *it != e) throw std::logic_error("update row missing in parent");
        o.map.push_back(uint32_t(it - pr));
      }
    }
  }
  for (uint32_t f = 0; f < F; ++f) o.child_start[f + 1] += o.child_start[f];
  o.pool = pool;
  o.minv_pool = minv;
  o.vec_pool = vec;
  o.y_pool = yv;
  o.g_pool = gp;
  // element pairs: lower node blocks (pos(i) >= pos(j)) of every tet
  struct Item {
    uint64_t key;
    uint32_t code;
  };
  for (size_t b = 0; b < bodies.size(); ++b) {
    const Body &bd = bodies[b];
    const Symbolic &s = sym[b];
    std::vector<Item> items;
    items.reserve(size_t(bd.tets) * 10);
    for (uint32_t e = 0; e < bd.tets; ++e) {
      const uint32_t *t = &bd.tet[4 * size_t(e)];
      for (uint32_t a = 0; a < 4; ++a)
        for (uint32_t c = 0; c < 4; ++c) {
          const int32_t pi = s.pos[t[a]], pj = s.pos[t[c]];
          if (pi < 0 || pj < 0 || pi < pj) continue;
          const uint64_t key = (uint64_t(uint32_t(pj)) << 32) | uint32_t(pi);
          items.push_back({key, ((t0[b] + e) << 4) | (a << 2) | c});
        }
    }
    std::sort(items.begin(), items.end(), [](const Item &x, const Item &y) {
      return x.key < y.key || (x.key == y.key && x.code < y.code);
    });
    // diagonal pairs for free nodes that have no tet (cannot happen for
    // meshes built here, but keep the mass term for every free node)
    size_t i = 0;
    while (i < items.size()) {
      size_t j = i;
      while (j < items.size() && items[j].key == items[i].key) ++j;
      const uint32_t pj = uint32_t(items[i].key >> 32), pi = uint32_t(items[i].key & 0xffffffffu);
      const uint32_t lf = s.front_of[pj];
      const FrontSym &fs = s.fronts[lf];
      const uint32_t f = fbase[b] + lf;
      const uint32_t col = pj - s.rows[fs.row0];
      const uint32_t *rr = &s.rows[fs.row0];
      const uint32_t *it = std::lower_bound(rr, rr + fs.nrow, pi);
      if (it == rr + fs.nrow || *it != pi) throw std::logic_error("pair row missing in front");
      const uint32_t row = uint32_t(it - rr);
      const uint64_t ld = o.m[f];
      o.pair_target.push_back(o.off[f] + uint64_t(3 * col) * ld + 3 * row);
      o.pair_ld.push_back(o.m[f]);
      o.pair_node.push_back(pi == pj ? n0[b] + s.perm[pi] : ~0u);
      o.pair_body.push_back(uint32_t(b));
      o.pair_start.push_back(uint32_t(o.contrib.size()));
      for (size_t q = i; q < j; ++q) o.contrib.push_back(items[q].code);
      i = j;
    }
  }
  o.pair_start.push_back(uint32_t(o.contrib.size()));
  // fronts per level
  const uint32_t L = o.levels;
  std::vector<std::vector<uint32_t>> by_level(L);
  for (uint32_t f = 0; f < F; ++f) by_level[o.level[f]].push_back(f);
  o.lvl_fronts_start.assign(1, 0);
  for (uint32_t l = 0; l < L; ++l) {
    o.lvl_fronts.insert(o.lvl_fronts.end(), by_level[l].begin(), by_level[l].end());
    o.lvl_fronts_start.push_back(uint32_t(o.lvl_fronts.size()));
  }
  // extend-add rounds: round r adds the r-th child of every front of the level
  o.ea_start.assign(1, 0);
  o.ea_rounds.assign(L, 0);
  for (uint32_t l = 0; l < L; ++l) {
    uint32_t rounds = 0;
    for (uint32_t f : by_level[l]) rounds = std::max(rounds, o.child_start[f + 1] - o.child_start[f]);
    o.ea_rounds[l] = rounds;
    for (uint32_t r = 0; r < rounds; ++r) {
      for (uint32_t f : by_level[l]) {
        if (o.child_start[f] + r >= o.child_start[f + 1]) continue;
        const uint32_t c = o.children[o.child_start[f] + r];
        const uint32_t u = o.m[c] - o.k[c];
        const uint32_t nt = tiles(u, FNB);
        for (uint32_t ti = 0; ti < nt; ++ti)
          for (uint32_t tj = 0; tj <= ti; ++tj)
            o.ea_tasks.push_back(uint64_t(c) | (uint64_t(ti) << 32) | (uint64_t(tj) << 48));
      }
      o.ea_start.push_back(uint32_t(o.ea_tasks.size()));
    }
  }
  // factorization launches: per level, per panel p
  o.fac_panels.assign(L, 0);
  o.potrf_start.assign(1, 0);
  o.trsm_start.assign(1, 0);
  o.syrk_start.assign(1, 0);
  for (uint32_t l = 0; l < L; ++l) {
    uint32_t panels = 0;
    for (uint32_t f : by_level[l]) panels = std::max(panels, tiles(o.k[f], FNB));
    o.fac_panels[l] = panels;
    for (uint32_t p = 0; p < panels; ++p) {
      for (uint32_t f : by_level[l]) {
        if (p * FNB >= o.k[f]) continue;
        o.potrf_tasks.push_back(f);
        // trailing rows start right after the panel: S = min(64 (p+1), k)
        const uint32_t S = std::min((p + 1) * FNB, o.k[f]);
        const uint32_t mt = tiles(o.m[f] - S, FNB);
        for (uint32_t rt = 0; rt < mt; ++rt) o.trsm_tasks.push_back(uint64_t(f) | (uint64_t(rt) << 32));
        for (uint32_t I = 0; I < mt; ++I)
          for (uint32_t J = 0; J <= I; ++J)
            o.syrk_tasks.push_back(uint64_t(f) | (uint64_t(I) << 32) | (uint64_t(J) << 48));
      }
      o.potrf_start.push_back(uint32_t(o.potrf_tasks.size()));
      o.trsm_start.push_back(uint32_t(o.trsm_tasks.size()));
      o.syrk_start.push_back(uint32_t(o.syrk_tasks.size()));
    }
  }
  // triangular inverse of L11: row block i, column blocks j < i
  uint32_t maxnb = 0;
  for (uint32_t f = 0; f < F; ++f) maxnb = std::max(maxnb, tiles(o.k[f], FNB));
  o.trtri_start.assign(1, 0);
  for (uint32_t i = 1; i < maxnb; ++i) {
    for (uint32_t f = 0; f < F; ++f) {
      if (i >= tiles(o.k[f], FNB)) continue;
      for (uint32_t j = 0; j < i; ++j)
        o.trtri_tasks.push_back(uint64_t(f) | (uint64_t(i) << 32) | (uint64_t(j) << 48));
    }
    o.trtri_start.push_back(uint32_t(o.trtri_tasks.size()));
  }
  // G packing: W21 tiles (rows of L21 x column blocks) and lower M11 tiles
  for (uint32_t f = 0; f < F; ++f) {
    const uint32_t kt = tiles(o.k[f], FNB), ut = tiles(o.m[f] - o.k[f], FNB);
    for (uint32_t I = 0; I < ut; ++I)
      for (uint32_t J = 0; J < kt; ++J) o.w21_tasks.push_back(uint64_t(f) | (uint64_t(I) << 32) | (uint64_t(J) << 48));
    for (uint32_t I = 0; I < kt; ++I)
      for (uint32_t J = 0; J <= I; ++J) o.pack_tasks.push_back(uint64_t(f) | (uint64_t(I) << 32) | (uint64_t(J) << 48));
  }
  // pull lists of the forward gather: child update rows in child order
  {
    std::vector<std::vector<uint64_t>> lists(o.rows.size());
    for (uint32_t f = 0; f < F; ++f)
      for (uint32_t ci = o.child_start[f]; ci < o.child_start[f + 1]; ++ci) {
        const uint32_t c = o.children[ci];
        const uint32_t kn = o.k[c] / 3, un = (o.m[c] - o.k[c]) / 3;
        for (uint32_t q = 0; q < un; ++q)
          lists[o.rows_off[f] + o.map[o.map_off[c] + q]].push_back(o.vec_off[c] + 3ull * (kn + q));
      }
    o.src_start.assign(1, 0);
    for (const auto &l : lists) {
      o.src.insert(o.src.end(), l.begin(), l.end());
      o.src_start.push_back(uint32_t(o.src.size()));
    }
  }
  // zeroing before a (partial) factorization: front storage m*m, M11 k*k
  for (uint32_t f = 0; f < o.fronts; ++f) {
    const uint64_t mm = uint64_t(o.m[f]) * o.m[f], kk = uint64_t(o.k[f]) * o.k[f];
    for (uint64_t c = 0; c * ZCHUNK < mm; ++c) o.zero_tasks.push_back(uint64_t(f) | (c << 32));
    for (uint64_t c = 0; c * ZCHUNK < kk; ++c) o.zero_tasks.push_back(uint64_t(f) | (c << 32) | (1ull << 48));
  }
  // solve: per level, gather chunks, forward row groups, x2 chunks, backward
  // column groups
  o.fwd_start.assign(1, 0);
  o.bwd_start.assign(1, 0);
  o.fwl_start.assign(1, 0);
  o.bwl_start.assign(1, 0);
  o.fwr_start.assign(1, 0);
  o.bwr_start.assign(1, 0);
  o.fws_start.assign(1, 0);
  o.bws_start.assign(1, 0);
  o.gat_all_start.assign(1, 0);
  o.gat_big_start.assign(1, 0);
  o.xg_all_start.assign(1, 0);
  o.xg_big_start.assign(1, 0);
  o.plan.assign(2 * size_t(L), 0);
  for (uint32_t l = 0; l < L; ++l) {
    for (uint32_t f : by_level[l]) {
      // rows kernel: 32 rows (columns) per workgroup, every front; lanes
      // kernel: one row (column) per lane, 256 per workgroup, for fronts with
      // few pivots, the rest of the level again with the rows kernel
      const bool fl = o.k[f] <= uint32_t(KLANES), bl = fl && o.m[f] <= uint32_t(MLANES);
      const uint64_t gs = ROWS_GROUP;
      for (uint32_t g = 0; g < tiles(o.m[f], uint32_t(gs)); ++g) {
        const uint64_t task = uint64_t(f) | (uint64_t(g) << 32) | (gs << 48);
        o.fwd_tasks.push_back(task);
        if (!fl) o.fwr_tasks.push_back(task);
      }
      for (uint32_t g = 0; g < tiles(o.k[f], uint32_t(gs)); ++g) {
        const uint64_t task = uint64_t(f) | (uint64_t(g) << 32) | (gs << 48);
        o.bwd_tasks.push_back(task);
        if (!bl) o.bwr_tasks.push_back(task);
      }
      for (uint32_t g = 0; g < tiles(o.m[f], GCHUNK); ++g) {
        o.gat_all.push_back(uint64_t(f) | (uint64_t(g) << 32));
        if (!fl) o.gat_big.push_back(uint64_t(f) | (uint64_t(g) << 32));
      }
      for (uint32_t g = 0; g < tiles(o.m[f] - o.k[f], GCHUNK); ++g) {
        o.xg_all.push_back(uint64_t(f) | (uint64_t(g) << 32));
        if (!bl) o.xg_big.push_back(uint64_t(f) | (uint64_t(g) << 32));
      }
      if (fl) {
        for (uint32_t g = 0; g < tiles(o.m[f], 256u); ++g) o.fwl_tasks.push_back(uint64_t(f) | (uint64_t(g) << 32));
        for (uint32_t g = 0; g < tiles(o.m[f], 64u); ++g) o.fws_tasks.push_back(uint64_t(f) | (uint64_t(g) << 32));
      }
      if (bl) {
        for (uint32_t g = 0; g < tiles(o.k[f], 256u); ++g) o.bwl_tasks.push_back(uint64_t(f) | (uint64_t(g) << 32));
        for (uint32_t g = 0; g < tiles(o.k[f], 64u); ++g) o.bws_tasks.push_back(uint64_t(f) | (uint64_t(g) << 32));
      }
    }
    o.fwd_start.push_back(uint32_t(o.fwd_tasks.size()));
    o.bwd_start.push_back(uint32_t(o.bwd_tasks.size()));
    o.fwl_start.push_back(uint32_t(o.fwl_tasks.size()));
    o.bwl_start.push_back(uint32_t(o.bwl_tasks.size()));
    o.fwr_start.push_back(uint32_t(o.fwr_tasks.size()));
    o.bwr_start.push_back(uint32_t(o.bwr_tasks.size()));
    o.fws_start.push_back(uint32_t(o.fws_tasks.size()));
    o.bws_start.push_back(uint32_t(o.bws_tasks.size()));
    o.gat_all_start.push_back(uint32_t(o.gat_all.size()));
    o.gat_big_start.push_back(uint32_t(o.gat_big.size()));
    o.xg_all_start.push_back(uint32_t(o.xg_all.size()));
    o.xg_big_start.push_back(uint32_t(o.xg_big.size()));
    o.plan[2 * l] = o.fwl_start[l + 1] - o.fwl_start[l] >= LANES_MIN_GROUPS ? 1 : 0;
    o.plan[2 * l + 1] = o.bwl_start[l + 1] - o.bwl_start[l] >= LANES_MIN_GROUPS ? 1 : 0;
  }
}

}  // namespace tmi
