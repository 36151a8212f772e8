// Host side of the strainstep hex lattices in libtalus_mi against the hex
// oracle, in one process on the CPU: tearCells (LatticeBody::tear_cells, with
// per-node cell lists and a hashed crack lookup) equals the oracle's
// restatement (linear scans) node for node on random deformed lattices with
// holes and anchored faces (surface nodes included), and lowestAngular (band L D L^T inverse
// iteration) equals the oracle's bit for bit on held and free beams.
// Built by tests/test_lattice_host_cpu.py with g++ against the built
// libraries; the oracle is the checker only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../talus_amd/csrc/hexlattice.h"
extern "C" {
#include "../../oracle/hex_oracle.h"
}

namespace {

int check_tears() {
  int bad = 0;
  for (int seed = 0; seed < 40; ++seed) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> N(0, 1);
    talus_mi_material m{};
    m.density = 20.0f; m.youngs_modulus = 5e3f; m.poissons_ratio = 0.3f; m.friction = 0.7f;
    m.strength = INFINITY; m.yield_stress = INFINITY; m.damping_ratio = INFINITY; m.tear_strain = 0.4f;
    m.plastic_max = INFINITY; m.crush_stress = INFINITY; m.crush_limit = 1.0f; m.hysteresis = 0.0f;
    const int nx = 6 + seed % 5, ny = 3 + seed % 3, nz = 2 + seed % 2;
    const double cell = 0.1;
    std::vector<float> rest; std::vector<float> mass; std::vector<uint8_t> anch; std::vector<uint32_t> hex;
    std::vector<float> vol; std::vector<uint32_t> lane;
    auto id = [&](int i, int j, int k) { return uint32_t(i + (nx + 1) * (j + (ny + 1) * k)); };
    for (int k = 0; k <= nz; ++k) for (int j = 0; j <= ny; ++j) for (int i = 0; i <= nx; ++i) {
      rest.push_back(float(i * cell)); rest.push_back(float(j * cell)); rest.push_back(float(k * cell));
      mass.push_back(0.0f); anch.push_back(i == 0 && (seed & 1));
    }
    for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
      if (seed % 3 == 2 && (i * 7 + j * 3 + k) % 11 == 0) continue;  // holes
      for (int c = 0; c < 8; ++c) hex.push_back(id(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)));
      vol.push_back(float(cell * cell * cell)); lane.push_back(0);
    }
    const uint32_t H = uint32_t(vol.size());
    for (uint32_t c = 0; c < H; ++c) for (int q = 0; q < 8; ++q) mass[hex[8 * c + q]] += 20.0f * vol[c] * 0.125f;
    // holes leave nodes outside every hex: give them a mass of their own
    for (auto &w : mass) if (w == 0.0f) w = 1.0f;
    talus_mi_hex_mesh mesh{};
    mesh.nodes = uint32_t(mass.size()); mesh.hexes = H;
    mesh.rest_positions = rest.data(); mesh.node_mass = mass.data(); mesh.anchored = anch.data();
    mesh.hex_nodes = hex.data(); mesh.volume = vol.data(); mesh.hex_material = lane.data();
    mesh.voxel_size[0] = mesh.voxel_size[1] = mesh.voxel_size[2] = float(cell);
    const double zero[3] = {0, 0, 0};
    tmi::LatticeBody lb;
    if (tmi::make_lattice(lb, &mesh, &m, 1, zero, zero) != 0) { printf("make failed %d\n", seed); return 1; }
    orc_hex_mesh om{};
    std::vector<double> rd(rest.begin(), rest.end());
    om.nodes = mesh.nodes; om.hexes = H; om.rest = rd.data(); om.mass = mass.data(); om.anchored = anch.data();
    om.hex = hex.data(); om.volume = vol.data(); om.lane = lane.data();
    om.voxel[0] = om.voxel[1] = om.voxel[2] = float(cell);
    orc_material omat; static_assert(sizeof omat == sizeof m, "material layout");
    std::memcpy(&omat, &m, sizeof m);
    orc_lattice_body *ob = orc_lattice_create(&om, &omat, 1, zero, zero);
    double A[9];
    for (int q = 0; q < 9; ++q) A[q] = (q % 4 == 0 ? 1.0 : 0.0) + 0.3 * N(rng);
    for (uint32_t i = 0; i < lb.nodes; ++i) {
      double p[3] = {lb.rest[3 * i], lb.rest[3 * i + 1], lb.rest[3 * i + 2]};
      for (int r = 0; r < 3; ++r) {
        lb.x[3 * i + r] = A[3 * r] * p[0] + A[3 * r + 1] * p[1] + A[3 * r + 2] * p[2] + 0.03 * std::sin(9.0 * p[(r + 1) % 3]);
        lb.v[3 * i + r] = N(rng);
      }
    }
    orc_lattice_set_state(ob, lb.x.data(), lb.v.data());
    std::vector<uint32_t> cells;
    for (int q = 0; q < 10; ++q) cells.push_back(uint32_t(rng() % H));
    const uint32_t a = lb.tear_cells(cells.data(), uint32_t(cells.size()));
    const uint32_t b = orc_lattice_tear(ob, cells.data(), uint32_t(cells.size()));
    const uint32_t on = orc_lattice_topology(ob, nullptr, nullptr, nullptr);
    std::vector<uint32_t> ohx(8 * H); std::vector<float> om2(on); std::vector<uint8_t> ocr(H);
    orc_lattice_topology(ob, ohx.data(), om2.data(), ocr.data());
    std::vector<double> ox(3 * on), ov(3 * on);
    orc_lattice_get_state(ob, ox.data(), ov.data());
    std::vector<uint8_t> lcr = lb.crack.empty() ? std::vector<uint8_t>(H, 0) : lb.crack;
    std::vector<uint32_t> osurf(on), lsurf;
    osurf.resize(orc_lattice_surface(ob, osurf.data()));
    lb.ensure_surface();
    for (uint32_t i = 0; i < lb.nodes; ++i)
      if (lb.surface[i]) lsurf.push_back(i);
    const bool same = a == b && on == lb.nodes && ohx == lb.hex && om2 == lb.mass && ocr == lcr && ox == lb.x &&
                      ov == lb.v && orc_lattice_pieces(ob) == lb.pieces() && osurf == lsurf;
    printf("seed %2d cells %3u copies %3u/%3u nodes %4u pieces %u/%u %s\n", seed, H, a, b, lb.nodes, lb.pieces(),
           orc_lattice_pieces(ob), same ? "same" : "DIFFERENT");
    bad += !same;
    orc_lattice_free(ob);
  }
  printf("tear different %d of 40\n", bad);
  return bad;
}

int check_lowest() {
  int bad = 0;
  for (int kase = 0; kase < 6; ++kase) {
    talus_mi_material m{};
    m.density = 5.0f; m.youngs_modulus = kase & 1 ? 5e5f : 1e5f; m.poissons_ratio = kase < 2 ? 0.0f : 0.3f + 0.05f * kase;
    m.friction = 0.5f; m.strength = INFINITY; m.yield_stress = INFINITY; m.damping_ratio = 0.2f;
    m.tear_strain = INFINITY; m.plastic_max = INFINITY; m.crush_stress = INFINITY; m.crush_limit = 1.0f;
    const int nx = 10 + 2 * kase, ny = 2, nz = 2;
    const double cell = 0.2;
    std::vector<float> rest, mass, vol; std::vector<uint8_t> anch; std::vector<uint32_t> hex, lane;
    auto id = [&](int i, int j, int k) { return uint32_t(i + (nx + 1) * (j + (ny + 1) * k)); };
    for (int k = 0; k <= nz; ++k) for (int j = 0; j <= ny; ++j) for (int i = 0; i <= nx; ++i) {
      rest.push_back(float(i * cell)); rest.push_back(float(j * cell)); rest.push_back(float(k * cell));
      mass.push_back(0.0f); anch.push_back(kase % 3 != 2 && i == 0);
    }
    for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
      for (int c = 0; c < 8; ++c) hex.push_back(id(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)));
      vol.push_back(float(cell * cell * cell)); lane.push_back(0);
    }
    const uint32_t H = uint32_t(vol.size());
    for (uint32_t c = 0; c < H; ++c) for (int q = 0; q < 8; ++q) mass[hex[8 * c + q]] += 5.0f * vol[c] * 0.125f;
    talus_mi_hex_mesh mesh{};
    mesh.nodes = uint32_t(mass.size()); mesh.hexes = H;
    mesh.rest_positions = rest.data(); mesh.node_mass = mass.data(); mesh.anchored = anch.data();
    mesh.hex_nodes = hex.data(); mesh.volume = vol.data(); mesh.hex_material = lane.data();
    mesh.voxel_size[0] = mesh.voxel_size[1] = mesh.voxel_size[2] = float(cell);
    const double zero[3] = {0, 0, 0};
    tmi::LatticeBody lb;
    if (tmi::make_lattice(lb, &mesh, &m, 1, zero, zero) != 0) { printf("make failed\n"); return 1; }
    orc_hex_mesh om{};
    std::vector<double> rd(rest.begin(), rest.end());
    om.nodes = mesh.nodes; om.hexes = H; om.rest = rd.data(); om.mass = mass.data(); om.anchored = anch.data();
    om.hex = hex.data(); om.volume = vol.data(); om.lane = lane.data();
    om.voxel[0] = om.voxel[1] = om.voxel[2] = float(cell);
    orc_material omat; std::memcpy(&omat, &m, sizeof m);
    orc_lattice_body *ob = orc_lattice_create(&om, &omat, 1, zero, zero);
    const double a = lb.lowest_angular(1.0 / 240.0), b = orc_lattice_lowest_angular(ob);
    // a damped build at h = 1/960 against the oracle's first damped step
    const double step = 1.0 / 960.0;
    lb.build_elements(step);
    const double g[3] = {0.0, -9.81, 0.0};
    orc_lattice_set_substeps(ob, 16);
    orc_lattice_step(ob, 1.0 / 60.0, g, 4, 1);
    const bool same = a == b && lb.lowest == orc_lattice_lowest(ob);
    printf("case %d cells %u lowest %.15g / %.15g built %.15g / %.15g %s\n", kase, H, a, b, lb.lowest,
           orc_lattice_lowest(ob), same ? "same" : "DIFFERENT");
    bad += !same;
    orc_lattice_free(ob);
  }
  printf("lowest different %d of 6\n", bad);
  return bad;
}

}  // namespace

int main() { return (check_tears() != 0) | ((check_lowest() != 0) << 1); }
