// mesh.cpp -- host mesh building for libtalus_mi.
//
// talus_mi_build_lattice replaces deformable::buildMesh for tet elements
// (src/talus/deformable/solver/solver.cc:2185-2418) on structured lattices
// whose solid cells are completely filled: nodes are numbered on first touch
// in a z, y, x scan of solid cells with corners 0..7 at offsets
// (c&1, c>>1&1, c>>2&1) (solver.cc:76-78, 2217-2249); every cell is split
// into the six Kuhn tets (solver.cc:17-24) around the 0-7 diagonal, a tet
// with negative orientation has nodes 1 and 2 swapped (2342-2350), its
// volume is det/6 (2352), rest inverse inverse(Ds) (2358) and each corner
// receives rho V / 4 of mass (2360-2364); a node is anchored when it lies in
// a support volume (163-180, 2265-2268).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/talus_mi.h"
#include "host.h"

namespace tmi {

static const uint32_t KUHN[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7},
                                    {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};

const talus_mi_material &lane_material(const talus_mi_material *materials, uint32_t lanes,
                                       uint32_t lane) {
  static const talus_mi_material fallback = default_material();
  return lane < lanes ? materials[lane] : fallback;
}

talus_mi_material default_material() {
  talus_mi_material m;
  m.friction = 0.5f;
  m.restitution = 0.0f;
  m.density = 1000.0f;
  m.strength = INFINITY;
  m.youngs_modulus = 0.0f;
  m.poissons_ratio = 0.3f;
  m.stiffness_damping = 0.0f;
  m.yield_stress = INFINITY;
  m.hardening_modulus = 0.0f;
  m.elastic_model = 0u;
  m.hysteresis = 0.0f;
  m.damping_ratio = INFINITY;
  m.tear_strain = INFINITY;
  m.plastic_max = INFINITY;
  m.crush_stress = INFINITY;
  m.crush_limit = 1.0f;
  return m;
}

void precise_lame(float youngs, float poisson, double *lambda, double *mu) {
  double nu = double(poisson);
  if (!(nu > 0.0)) nu = 0.0;
  if (nu > 0.49) nu = 0.49;
  const double e = double(youngs);
  *lambda = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  *mu = e / (2.0 * (1.0 + nu));
}

// Rest edge matrix Ds (columns X_k - X_0) in FP64 and its inverse.
bool rest_frame(const double *rest, const uint32_t *tet, double *dminv, double *volume) {
  double ds[9];
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r)
      ds[r * 3 + k] = rest[3 * size_t(tet[k + 1]) + r] - rest[3 * size_t(tet[0]) + r];
  const double det = ds[0] * (ds[4] * ds[8] - ds[5] * ds[7]) -
                     ds[1] * (ds[3] * ds[8] - ds[5] * ds[6]) +
                     ds[2] * (ds[3] * ds[7] - ds[4] * ds[6]);
  *volume = det / 6.0;
  if (!(det > 0.0)) return false;
  const double s = 1.0 / det;
  dminv[0] = (ds[4] * ds[8] - ds[5] * ds[7]) * s;
  dminv[1] = (ds[2] * ds[7] - ds[1] * ds[8]) * s;
  dminv[2] = (ds[1] * ds[5] - ds[2] * ds[4]) * s;
  dminv[3] = (ds[5] * ds[6] - ds[3] * ds[8]) * s;
  dminv[4] = (ds[0] * ds[8] - ds[2] * ds[6]) * s;
  dminv[5] = (ds[2] * ds[3] - ds[0] * ds[5]) * s;
  dminv[6] = (ds[3] * ds[7] - ds[4] * ds[6]) * s;
  dminv[7] = (ds[1] * ds[6] - ds[0] * ds[7]) * s;
  dminv[8] = (ds[0] * ds[4] - ds[1] * ds[3]) * s;
  return true;
}

int build_lattice(const talus_mi_lattice *L, const talus_mi_material *materials, uint32_t lanes,
                  const talus_mi_support *supports, uint32_t nsupports, talus_mi_mesh *out) {
  std::memset(out, 0, sizeof *out);
  for (int k = 0; k < 3; ++k)
    if (L->resolution[k] < 2) return fail(TALUS_MI_ERR_ARGUMENT, "lattice resolution below 2");
  const uint32_t cx = L->resolution[0] - 1, cy = L->resolution[1] - 1, cz = L->resolution[2] - 1;
  const size_t rx = L->resolution[0], ry = L->resolution[1], rz = L->resolution[2];
  const size_t samples = rx * ry * rz;
  std::vector<uint32_t> node_of(samples, UINT32_MAX);
  std::vector<uint32_t> sample_of;
  size_t solid_cells = 0;
  auto sample_index = [&](uint32_t x, uint32_t y, uint32_t z) { return x + rx * (y + ry * size_t(z)); };
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        const size_t cell = x + size_t(cx) * (y + size_t(cy) * z);
        if (L->solid && !L->solid[cell]) continue;
        ++solid_cells;
        for (uint32_t c = 0; c < 8; ++c) {
          const size_t s = sample_index(x + (c & 1u), y + ((c >> 1) & 1u), z + ((c >> 2) & 1u));
          if (node_of[s] == UINT32_MAX) {
            node_of[s] = uint32_t(sample_of.size());
            sample_of.push_back(uint32_t(s));
          }
        }
      }
  const uint32_t nodes = uint32_t(sample_of.size());
  if (nodes == 0) return fail(TALUS_MI_ERR_ARGUMENT, "lattice has no solid cell");
  std::vector<float> restf(3 * size_t(nodes));
  std::vector<double> rest(3 * size_t(nodes));
  for (uint32_t n = 0; n < nodes; ++n) {
    const size_t s = sample_of[n];
    float p[3];
    if (L->sample_positions) {
      for (int k = 0; k < 3; ++k) p[k] = L->sample_positions[3 * s + k];
    } else {
      const uint32_t ax = uint32_t(s % rx), ay = uint32_t((s / rx) % ry), az = uint32_t(s / (rx * ry));
      p[0] = L->origin[0] + L->voxel_size[0] * float(ax);
      p[1] = L->origin[1] + L->voxel_size[1] * float(ay);
      p[2] = L->origin[2] + L->voxel_size[2] * float(az);
    }
    for (int k = 0; k < 3; ++k) {
      restf[3 * size_t(n) + k] = p[k];
      rest[3 * size_t(n) + k] = double(p[k]);
    }
  }
  std::vector<uint8_t> anchored(nodes, 0);
  for (uint32_t n = 0; n < nodes; ++n)
    for (uint32_t v = 0; v < nsupports; ++v) {
      const float *M = supports[v].world_to_support;
      bool inside = true;
      for (int r = 0; r < 3 && inside; ++r) {
        const float l = M[r * 4 + 0] * restf[3 * size_t(n)] + M[r * 4 + 1] * restf[3 * size_t(n) + 1] +
                        M[r * 4 + 2] * restf[3 * size_t(n) + 2] + M[r * 4 + 3];
        if (!(std::fabs(l) <= supports[v].half_extents[r])) inside = false;
      }
      if (inside) {
        anchored[n] = 1;
        break;
      }
    }
  std::vector<uint32_t> tets;
  std::vector<double> dminv, volume;
  std::vector<uint32_t> lane_of;
  std::vector<double> mass(nodes, 0.0);
  tets.reserve(24 * solid_cells);
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        const size_t cell = x + size_t(cx) * (y + size_t(cy) * z);
        if (L->solid && !L->solid[cell]) continue;
        uint32_t corners[8];
        for (uint32_t c = 0; c < 8; ++c)
          corners[c] = node_of[sample_index(x + (c & 1u), y + ((c >> 1) & 1u), z + ((c >> 2) & 1u))];
        const uint32_t lane = L->cell_lane ? L->cell_lane[cell] : 0u;
        const double density = double(lane_material(materials, lanes, lane).density);
        for (int k = 0; k < 6; ++k) {
          uint32_t tn[4] = {corners[KUHN[k][0]], corners[KUHN[k][1]], corners[KUHN[k][2]],
                            corners[KUHN[k][3]]};
          double dm[9], vol;
          if (!rest_frame(rest.data(), tn, dm, &vol)) {
            std::swap(tn[1], tn[2]);
            if (!rest_frame(rest.data(), tn, dm, &vol)) continue;
          }
          tets.insert(tets.end(), tn, tn + 4);
          dminv.insert(dminv.end(), dm, dm + 9);
          volume.push_back(vol);
          lane_of.push_back(lane);
          for (int c = 0; c < 4; ++c) mass[tn[c]] += density * vol * 0.25;
        }
      }
  const uint32_t ntets = uint32_t(volume.size());
  if (ntets == 0) return fail(TALUS_MI_ERR_ARGUMENT, "lattice produced no tet");
  out->nodes = nodes;
  out->tets = ntets;
  out->rest_positions = (float *)std::malloc(restf.size() * sizeof(float));
  out->node_mass = (float *)std::malloc(nodes * sizeof(float));
  out->anchored = (uint8_t *)std::malloc(nodes);
  out->tet_nodes = (uint32_t *)std::malloc(tets.size() * sizeof(uint32_t));
  out->rest_inverse = (float *)std::malloc(dminv.size() * sizeof(float));
  out->volume = (float *)std::malloc(ntets * sizeof(float));
  out->tet_material = (uint32_t *)std::malloc(ntets * sizeof(uint32_t));
  if (!out->rest_positions || !out->node_mass || !out->anchored || !out->tet_nodes ||
      !out->rest_inverse || !out->volume || !out->tet_material) {
    mesh_free(out);
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory building a mesh");
  }
  std::memcpy(out->rest_positions, restf.data(), restf.size() * sizeof(float));
  for (uint32_t n = 0; n < nodes; ++n) out->node_mass[n] = float(mass[n]);
  std::memcpy(out->anchored, anchored.data(), nodes);
  std::memcpy(out->tet_nodes, tets.data(), tets.size() * sizeof(uint32_t));
  for (size_t i = 0; i < dminv.size(); ++i) out->rest_inverse[i] = float(dminv[i]);
  for (uint32_t e = 0; e < ntets; ++e) out->volume[e] = float(volume[e]);
  std::memcpy(out->tet_material, lane_of.data(), ntets * sizeof(uint32_t));
  return TALUS_MI_OK;
}

void mesh_free(talus_mi_mesh *m) {
  if (!m) return;
  std::free(m->rest_positions);
  std::free(m->node_mass);
  std::free(m->anchored);
  std::free(m->tet_nodes);
  std::free(m->rest_inverse);
  std::free(m->volume);
  std::free(m->tet_material);
  std::memset(m, 0, sizeof *m);
}

}  // namespace tmi
