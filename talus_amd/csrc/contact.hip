// contact.hip -- findContacts on the GPU (contact.h).  Every float
// operation follows the reference's order with contraction off (this file is
// built with -ffp-contract=off and correctly rounded division and square
// root), so a contact is the reference's bit for bit; the shape's turn uses
// sin and cos evaluated in FP64 and rounded, i.e. correctly rounded, where
// the reference's C library returns within 0.56 ulp of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "contact.h"
#include "device.h"

namespace tmi {

namespace {

using namespace contact_math;


constexpr uint32_t NEIGHBOURS = 8u;          // earlier candidates listed per candidate
constexpr int SELECT_THREADS = 1024;
enum : uint8_t { UNDECIDED = 0, ACCEPTED = 1, REJECTED = 2 };

// ------------------------------------------------------------- kernels

struct ClusterData {
  float lo[3], hi[3], centre[3];
  float spread, speed, thickness;
};

struct ContactDev {
  uint32_t N, T, C, S, lanes, capacity;
  const uint32_t *tri, *lane, *cfaces;
  const uint8_t *anchored;
  const float *friction, *thickness;
  const double *x, *v;
  float *pos, *vel;
  ClusterData *clusters;
  uint8_t *skip, *flag, *state;
  Candidate *cand;
  uint32_t *order, *nbr, *nbn, *count, *kount;
  float *pts;  // candidates' mesh points in face order, 3 per candidate
  talus_mi_surface_contact *out;
  const DevShape *shapes;
  ShapePre *pre;  // per shape
  float dt, g[3], bias, threshold;
  float near2;  // the largest squared distance whose square root is <= apart
  int cull, clustered;
};

// the float state findContacts reads; velocityOf (contact.cc:624-628)
__global__ void k_contact_state(ContactDev d) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.N) return;
  const bool held = d.anchored[n] != 0;
  for (int k = 0; k < 3; ++k) {
    d.pos[3 * n + k] = float(d.x[3 * size_t(n) + k]);
    d.vel[3 * n + k] = held ? 0.0f : float(d.v[3 * size_t(n) + k]) + d.g[k] * d.dt;
  }
}

__device__ inline float red_min(float v) {
  for (int o = 16; o > 0; o >>= 1) v = mn(v, __shfl_xor(v, o, 32));
  return v;
}
__device__ inline float red_max(float v) {
  for (int o = 16; o > 0; o >>= 1) v = mx(v, __shfl_xor(v, o, 32));
  return v;
}

// the shapes' swept boxes, radii and starting poses, once per query
__global__ void k_contact_shapes(ContactDev d) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < d.S) d.pre[s] = shape_pre(d.shapes[s], d.dt);
}

// ClusterReach (contact.cc:640-703) and clusterFar (454-481) per shape: one
// 32-lane half-wavefront per cluster, one lane per triangle; the box, speed,
// skin and spread are minima and maxima, the same in any order
__global__ void __launch_bounds__(256) k_contact_clusters(ContactDev d) {
  const uint32_t lane = threadIdx.x & 31u;
  const uint32_t c = blockIdx.x * (blockDim.x / 32u) + threadIdx.x / 32u;
  if (c >= d.C) return;  // whole half-wavefronts leave together
  const uint32_t at = c * SURFACE_CLUSTER + lane;
  const bool active = at < d.T;
  Bounds start = empty_bounds(), reach = empty_bounds();
  float speed = 0.0f, thick = 0.0f;
  uint32_t face = 0;
  if (active) {
    face = d.cfaces[at];
    thick = d.thickness ? d.thickness[face] : 0.0f;
    for (int k = 0; k < 3; ++k) {
      const uint32_t node = d.tri[3 * size_t(face) + k];
      const F3 p = ld3(&d.pos[3 * size_t(node)]), v = ld3(&d.vel[3 * size_t(node)]);
      include(start, p);
      include(reach, p);
      include(reach, p + v * d.dt);
      speed = mx(speed, len(v));
    }
  }
  start.lo = mk(red_min(start.lo.x), red_min(start.lo.y), red_min(start.lo.z));
  start.hi = mk(red_max(start.hi.x), red_max(start.hi.y), red_max(start.hi.z));
  reach.lo = mk(red_min(reach.lo.x), red_min(reach.lo.y), red_min(reach.lo.z));
  reach.hi = mk(red_max(reach.hi.x), red_max(reach.hi.y), red_max(reach.hi.z));
  speed = red_max(speed);
  thick = red_max(mx(0.0f, thick));
  const F3 centre = (start.lo + start.hi) * 0.5f;
  float spread = 0.0f;
  if (active)
    for (int k = 0; k < 3; ++k)
      spread = mx(spread, len(centre - ld3(&d.pos[3 * size_t(d.tri[3 * size_t(face) + k])])));
  spread = red_max(spread);
  if (lane == 0) {
    ClusterData &o = d.clusters[c];
    o.lo[0] = reach.lo.x, o.lo[1] = reach.lo.y, o.lo[2] = reach.lo.z;
    o.hi[0] = reach.hi.x, o.hi[1] = reach.hi.y, o.hi[2] = reach.hi.z;
    o.centre[0] = centre.x, o.centre[1] = centre.y, o.centre[2] = centre.z;
    o.spread = spread, o.speed = speed, o.thickness = thick;
  }
  for (uint32_t s = lane; s < d.S; s += 32u) {
    const DevShape &sh = d.shapes[s];
    const ShapePre &pre = d.pre[s];
    const float sb = d.bias + thick, st = d.threshold + thick;
    bool far = !overlaps(reach, pre.swept, st);
    if (!far)
      far = out_of_reach(sh, distance_below(sh, pre.origin, centre, spread, sh.lipschitz), speed, d.dt, sb, st,
                         pre.radius);
    d.skip[size_t(s) * d.C + c] = far ? 1u : 0u;
  }
}

// one (shape, triangle) per thread, triangles in cluster order so that a
// wavefront probes one neighbourhood of the field
__global__ void __launch_bounds__(256) k_contact_faces(ContactDev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = blockIdx.y;
  if (i >= d.T) return;
  const uint32_t face = d.clustered ? d.cfaces[i] : i;
  uint8_t &flag = d.flag[size_t(s) * d.T + face];
  flag = 0;
  if (d.clustered && d.skip[size_t(s) * d.C + i / SURFACE_CLUSTER]) return;
  FaceQuery q;
  q.tri = d.tri, q.lane = d.lane, q.pos = d.pos, q.vel = d.vel, q.friction = d.friction;
  q.lanes = d.lanes, q.dt = d.dt, q.bias = d.bias, q.threshold = d.threshold, q.cull = d.cull;
  const float thick = d.thickness ? d.thickness[face] : 0.0f;
  if (face_contact(q, d.shapes[s], d.pre[s], int32_t(s), face, thick, &d.cand[size_t(s) * d.T + face])) flag = 1;
}

// exclusive prefix sum over the workgroup (SELECT_THREADS threads, 64-wide
// wavefronts); *total receives the sum
__device__ uint32_t block_scan(uint32_t value, uint32_t *total) {
  __shared__ uint32_t waves[SELECT_THREADS / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inclusive = value;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t up = __shfl_up(inclusive, o, 64);
    if (lane >= uint32_t(o)) inclusive += up;
  }
  if (lane == 63u) waves[wave] = inclusive;
  __syncthreads();
  uint32_t before = 0, sum = 0;
  for (uint32_t w = 0; w < SELECT_THREADS / 64; ++w) {
    if (w < wave) before += waves[w];
    sum += waves[w];
  }
  __syncthreads();
  *total = sum;
  return before + inclusive - value;
}

// the duplicate test (contact.cc:806-826): a candidate is kept when no
// earlier KEPT candidate of its shape (face order) lies within `apart` -- a
// sequential greedy rule, evaluated here in three steps: (1) each shape's
// candidates compacted in face order with their mesh points, (2) every
// candidate's earlier neighbours within `apart` found grid-wide, (3) the rule
// resolved in rounds (a candidate decides once all its earlier neighbours
// have; the first undecided one always can, so each round decides at least
// one) and the kept contacts written shape by shape

// (1) one workgroup per shape
__global__ void __launch_bounds__(SELECT_THREADS) k_contact_compact(ContactDev d) {
  const uint32_t s = blockIdx.x;
  const size_t at = size_t(s) * d.T;
  uint32_t K = 0;
  for (uint32_t base = 0; base < d.T; base += SELECT_THREADS) {
    const uint32_t f = base + threadIdx.x;
    const uint32_t fl = f < d.T ? d.flag[at + f] : 0u;
    uint32_t total;
    const uint32_t rank = block_scan(fl, &total);
    if (fl) {
      const size_t k = at + K + rank;
      d.order[k] = f;
      const float *p = d.cand[at + f].point;
      d.pts[3 * k] = p[0];
      d.pts[3 * k + 1] = p[1];
      d.pts[3 * k + 2] = p[2];
      d.nbn[k] = 0;
      d.state[k] = UNDECIDED;
    }
    K += total;
  }
  if (threadIdx.x == 0) d.kount[s] = K;
}

__device__ inline F3 point_of(const float *pts, size_t i) { return mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); }

// |p - q| <= apart, decided as dot(p - q, p - q) <= near2: the square root
// is correctly rounded and monotone, so this is the same decision
__device__ inline bool near(F3 p, F3 q, float near2) {
  const F3 e = p - q;
  return dot(e, e) <= near2;
}

// (2) all earlier neighbours: the lower triangle of 256 x 256 tiles of
// (candidate, earlier candidate) pairs, grid-stride over the tiles, the
// earlier tile's points staged in LDS; a pair found appends to the
// candidate's list (the first NEIGHBOURS listed, the count kept)
__global__ void __launch_bounds__(256) k_contact_neighbours(ContactDev d) {
  __shared__ float tile[3][256];
  const uint32_t s = blockIdx.y;
  const uint32_t K = d.kount[s];
  const uint64_t nt = (uint64_t(K) + 255u) / 256u;
  const uint64_t tiles = nt * (nt + 1) / 2;
  const size_t at = size_t(s) * d.T;
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    uint64_t kt = uint64_t((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
    while (kt * (kt + 1) / 2 > t) --kt;
    while ((kt + 1) * (kt + 2) / 2 <= t) ++kt;
    const uint32_t jt = uint32_t(t - kt * (kt + 1) / 2);
    const uint32_t j0 = jt * 256u, jn = min(256u, K - j0);
    __syncthreads();
    if (threadIdx.x < jn) {
      const F3 q = point_of(d.pts, at + j0 + threadIdx.x);
      tile[0][threadIdx.x] = q.x;
      tile[1][threadIdx.x] = q.y;
      tile[2][threadIdx.x] = q.z;
    }
    __syncthreads();
    const uint32_t k = uint32_t(kt) * 256u + threadIdx.x;
    if (k >= K) continue;
    const F3 me = point_of(d.pts, at + k);
    const uint32_t end = uint32_t(kt) == jt ? threadIdx.x : jn;
    for (uint32_t i = 0; i < end; ++i) {
      if (!near(mk(tile[0][i], tile[1][i], tile[2][i]), me, d.near2)) continue;
      const uint32_t slot = atomicAdd(&d.nbn[at + k], 1u);
      if (slot < NEIGHBOURS) d.nbr[(at + k) * NEIGHBOURS + slot] = j0 + i;
    }
  }
}

// (3) one workgroup for all shapes, in order
__global__ void __launch_bounds__(SELECT_THREADS) k_contact_resolve(ContactDev d) {
  __shared__ uint32_t undecided;
  uint32_t written = 0;
  for (uint32_t s = 0; s < d.S; ++s) {
    const size_t at = size_t(s) * d.T;
    const uint32_t K = d.kount[s];
    volatile uint8_t *state = d.state + at;
    for (;;) {
      if (threadIdx.x == 0) undecided = 0;
      __syncthreads();
      for (uint32_t k = threadIdx.x; k < K; k += SELECT_THREADS) {
        if (state[k] != UNDECIDED) continue;
        const uint32_t n = d.nbn[at + k];
        bool kept_before = false, open = false;
        if (n <= NEIGHBOURS) {
          for (uint32_t i = 0; i < n; ++i) {
            const uint8_t o = state[d.nbr[(at + k) * NEIGHBOURS + i]];
            kept_before |= o == ACCEPTED;
            open |= o == UNDECIDED;
          }
        } else {
          const F3 me = point_of(d.pts, at + k);
          for (uint32_t j = 0; j < k; ++j)
            if (near(point_of(d.pts, at + j), me, d.near2)) {
              const uint8_t o = state[j];
              kept_before |= o == ACCEPTED;
              open |= o == UNDECIDED;
            }
        }
        if (kept_before)
          state[k] = REJECTED;
        else if (!open)
          state[k] = ACCEPTED;
        else
          atomicAdd(&undecided, 1u);
      }
      __syncthreads();
      const uint32_t left = undecided;
      __syncthreads();
      if (left == 0) break;
    }
    uint32_t kept = 0;
    for (uint32_t base = 0; base < K; base += SELECT_THREADS) {
      const uint32_t k = base + threadIdx.x;
      const uint32_t acc = k < K && state[k] == ACCEPTED ? 1u : 0u;
      uint32_t total;
      const uint32_t rank = block_scan(acc, &total);
      const uint32_t o = written + kept + rank;
      if (acc && o < d.capacity) d.out[o] = d.cand[at + d.order[at + k]].c;
      kept += total;
    }
    written += kept;
    __syncthreads();
  }
  if (threadIdx.x == 0) *d.count = written;
}

// cellLipschitz (sampledsdf.cc:61-93) of every cell; the maximum as the
// bits of a non-negative float (ordered like unsigned integers)
__global__ void k_field_lipschitz(FieldView f, uint32_t *worst) {
  const uint32_t cx = f.res[0] - 1u, cy = f.res[1] - 1u, cz = f.res[2] - 1u;
  const size_t cells = size_t(cx) * cy * cz;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cells) return;
  Cell cell;
  cell.base[0] = uint32_t(i % cx);
  cell.base[1] = uint32_t((i / cx) % cy);
  cell.base[2] = uint32_t(i / (size_t(cx) * cy));
  float c[8];
  corners(f, cell, c);
  F3 s = mk(0.0f, 0.0f, 0.0f);
  for (uint32_t k = 0; k < 8u; ++k) {
    if (!(k & 1u)) s.x = mx(s.x, fabsf(c[k | 1u] - c[k]));
    if (!(k & 2u)) s.y = mx(s.y, fabsf(c[k | 2u] - c[k]));
    if (!(k & 4u)) s.z = mx(s.z, fabsf(c[k | 4u] - c[k]));
  }
  const float l = len(mk(s.x / f.voxel[0], s.y / f.voxel[1], s.z / f.voxel[2]));
  atomicMax(worst, __float_as_uint(l));
}

template <class T>
T *alloc(size_t n, std::vector<void *> &owned) {
  T *ptr = nullptr;
  HIPCHECK(hipMalloc(&ptr, std::max<size_t>(n, 1) * sizeof(T)));
  owned.push_back(ptr);
  return ptr;
}

template <class T>
T *upload(const std::vector<T> &h, std::vector<void *> &owned) {
  T *ptr = alloc<T>(h.size(), owned);
  if (!h.empty()) HIPCHECK(hipMemcpy(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return ptr;
}

void release(void *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

// ------------------------------------------------------------ DeviceField

DeviceField::~DeviceField() {
  if (dist_) (void)hipFree(dist_);
  if (mat_) (void)hipFree(mat_);
}

void DeviceField::set(const talus_mi_sdf &field, hipStream_t stream) {
  for (int k = 0; k < 3; ++k)
    if (field.resolution[k] < 2u) throw std::invalid_argument("a sampled field needs two samples per axis");
  if (!field.distances) throw std::invalid_argument("a sampled field needs its distances");
  const size_t n = size_t(field.resolution[0]) * field.resolution[1] * field.resolution[2];
  if (n > capacity_) {
    if (dist_) HIPCHECK(hipFree(dist_));
    if (mat_) HIPCHECK(hipFree(mat_));
    dist_ = nullptr, mat_ = nullptr, capacity_ = 0;
    HIPCHECK(hipMalloc(&dist_, n * sizeof(float)));
    HIPCHECK(hipMalloc(&mat_, n));
    capacity_ = n;
  }
  HIPCHECK(hipMemcpyAsync(dist_, field.distances, n * sizeof(float), hipMemcpyHostToDevice, stream));
  if (field.materials)
    HIPCHECK(hipMemcpyAsync(mat_, field.materials, n, hipMemcpyHostToDevice, stream));
  else
    HIPCHECK(hipMemsetAsync(mat_, 0, n, stream));
  for (int k = 0; k < 3; ++k) {
    view_.origin[k] = field.origin[k];
    view_.voxel[k] = field.voxel_size[k];
    view_.res[k] = field.resolution[k];
  }
  view_.dist = dist_;
  view_.mat = mat_;
  uint32_t *worst = nullptr;
  HIPCHECK(hipMalloc(&worst, sizeof(uint32_t)));
  uint32_t bits = 0;
  try {
    HIPCHECK(hipMemsetAsync(worst, 0, sizeof(uint32_t), stream));
    const size_t cells = size_t(view_.res[0] - 1u) * (view_.res[1] - 1u) * (view_.res[2] - 1u);
    hipLaunchKernelGGL(k_field_lipschitz, dim3(uint32_t((cells + 255) / 256)), dim3(256), 0, stream, view_, worst);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(&bits, worst, sizeof bits, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  } catch (...) {
    (void)hipFree(worst);
    throw;
  }
  HIPCHECK(hipFree(worst));
  float g;
  std::memcpy(&g, &bits, sizeof g);
  // probeLipschitz (sampledsdf.cc:424-430) of a finished field
  lipschitz_ = std::sqrt(g * g + 1.0f);
}

// ---------------------------------------------------------- ContactFinder

ContactFinder::ContactFinder(const Body &b) {
  SurfaceMesh sm;
  mesh_surface(b.nodes, b.tets, b.tet.data(), b.lane.data(), b.rest.data(), sm);
  N_ = b.nodes;
  T_ = uint32_t(sm.triangle_lane.size());
  C_ = sm.cluster_starts.empty() ? 0u : uint32_t(sm.cluster_starts.size() - 1);
  for (uint32_t c = 0; c <= C_ && T_; ++c)
    if (sm.cluster_starts[c] != std::min(c * SURFACE_CLUSTER, T_))
      throw std::logic_error("surface clusters are not 32 triangles each");
  lanes_ = uint32_t(b.materials.size());
  std::vector<float> friction(lanes_);
  for (uint32_t l = 0; l < lanes_; ++l) friction[l] = b.materials[l].friction;
  try {
    tri_ = upload(sm.triangles, owned_);
    lane_ = upload(sm.triangle_lane, owned_);
    cfaces_ = upload(sm.cluster_faces, owned_);
    anchored_ = upload(b.anchored, owned_);
    friction_ = upload(friction, owned_);
    thickness_ = alloc<float>(T_, owned_);
    pos_ = alloc<float>(3 * size_t(N_), owned_);
    vel_ = alloc<float>(3 * size_t(N_), owned_);
    clusters_ = alloc<ClusterData>(C_, owned_);
    count_ = alloc<uint32_t>(1, owned_);
  } catch (...) {
    for (void *p : owned_) (void)hipFree(p);
    throw;
  }
}

ContactFinder::~ContactFinder() {
  for (void *p : owned_) (void)hipFree(p);
  for (void *p : per_shape_) (void)hipFree(p);
  release(out_);
}

void ContactFinder::reserve(uint32_t shapes, uint32_t capacity) {
  if (shapes > shapes_cap_) {
    for (void *p : per_shape_) HIPCHECK(hipFree(p));
    per_shape_.clear();
    shapes_cap_ = 0;
    const size_t ST = std::max<size_t>(size_t(shapes) * T_, 1);
    dshapes_ = alloc<DevShape>(shapes, per_shape_);
    pre_ = alloc<ShapePre>(shapes, per_shape_);
    cand_ = alloc<Candidate>(ST, per_shape_);
    skip_ = alloc<uint8_t>(std::max<size_t>(size_t(shapes) * C_, 1), per_shape_);
    flag_ = alloc<uint8_t>(ST, per_shape_);
    state_ = alloc<uint8_t>(ST, per_shape_);
    order_ = alloc<uint32_t>(ST, per_shape_);
    nbn_ = alloc<uint32_t>(ST, per_shape_);
    nbr_ = alloc<uint32_t>(ST * NEIGHBOURS, per_shape_);
    pts_ = alloc<float>(3 * ST, per_shape_);
    kount_ = alloc<uint32_t>(shapes, per_shape_);
    shapes_cap_ = shapes;
  }
  if (capacity > out_cap_) {
    release(out_);
    out_cap_ = 0;
    HIPCHECK(hipMalloc(&out_, capacity * sizeof(talus_mi_surface_contact)));
    out_cap_ = capacity;
  }
}

uint32_t ContactFinder::find(const double *x, const double *v, hipStream_t stream, const talus_mi_shape *shapes,
                             const DeviceField *const *fields, uint32_t count, const talus_mi_contact_query &q,
                             const float *thickness, talus_mi_surface_contact *out, uint32_t capacity) {
  if (!count || !T_) return 0;
  reserve(count, std::max(capacity, 1u));
  std::vector<DevShape> host(count);
  for (uint32_t s = 0; s < count; ++s) {
    DevShape &d = host[s];
    const talus_mi_shape &in = shapes[s];
    d.f = fields[s]->view();
    for (int k = 0; k < 3; ++k) {
      d.center[k] = in.center_of_mass[k];
      d.offset[k] = in.field_offset[k];
      d.lin[k] = in.linear_velocity[k];
      d.ang[k] = in.angular_velocity[k];
    }
    for (int k = 0; k < 4; ++k) d.rot[k] = in.rotation[k];
    d.lipschitz = q.cull ? fields[s]->lipschitz() : 0.0f;
  }
  HIPCHECK(hipMemcpyAsync(dshapes_, host.data(), count * sizeof(DevShape), hipMemcpyHostToDevice, stream));
  if (thickness)
    HIPCHECK(hipMemcpyAsync(thickness_, thickness, T_ * sizeof(float), hipMemcpyHostToDevice, stream));
  ContactDev d{};
  d.N = N_, d.T = T_, d.C = C_, d.S = count, d.lanes = lanes_, d.capacity = capacity;
  d.tri = tri_, d.lane = lane_, d.cfaces = cfaces_, d.anchored = anchored_;
  d.friction = friction_, d.thickness = thickness ? thickness_ : nullptr;
  d.x = x, d.v = v, d.pos = pos_, d.vel = vel_;
  d.clusters = static_cast<ClusterData *>(clusters_);
  d.skip = skip_, d.flag = flag_, d.state = state_;
  d.cand = cand_;
  d.order = order_, d.nbr = nbr_, d.nbn = nbn_, d.count = count_, d.kount = kount_, d.pts = pts_;
  d.out = static_cast<talus_mi_surface_contact *>(out_);
  d.shapes = dshapes_;
  d.pre = pre_;
  d.dt = q.dt;
  for (int k = 0; k < 3; ++k) d.g[k] = q.gravity[k];
  d.bias = q.bias, d.threshold = q.threshold;
  // duplicates lie within apart = 1e-2 of the largest voxel side
  // (contact.cc:630); sqrtf is correctly rounded and monotone, so
  // sqrtf(x) <= apart exactly when x <= near2
  const float apart = 1.0e-2f * std::max(std::max(q.voxel_size[0], q.voxel_size[1]), q.voxel_size[2]);
  float near2 = apart * apart;
  if (!(apart >= 0.0f) || !std::isfinite(apart)) {
    near2 = apart >= 0.0f ? INFINITY : -1.0f;
  } else {
    while (near2 > 0.0f && std::sqrt(near2) > apart) near2 = std::nextafter(near2, 0.0f);
    while (std::sqrt(std::nextafter(near2, INFINITY)) <= apart) near2 = std::nextafter(near2, INFINITY);
  }
  d.near2 = near2;
  d.cull = q.cull ? 1 : 0;
  d.clustered = q.cull && C_ ? 1 : 0;
  hipLaunchKernelGGL(k_contact_state, dim3((N_ + 255) / 256), dim3(256), 0, stream, d);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_contact_shapes, dim3((count + 63) / 64), dim3(64), 0, stream, d);
  HIPCHECK(hipGetLastError());
  if (d.clustered) {
    hipLaunchKernelGGL(k_contact_clusters, dim3((C_ + 7) / 8), dim3(256), 0, stream, d);
    HIPCHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_contact_faces, dim3((T_ + 255) / 256, count), dim3(256), 0, stream, d);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_contact_compact, dim3(count), dim3(SELECT_THREADS), 0, stream, d);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_contact_neighbours, dim3(256, count), dim3(256), 0, stream, d);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_contact_resolve, dim3(1), dim3(SELECT_THREADS), 0, stream, d);
  HIPCHECK(hipGetLastError());
  uint32_t found = 0;
  HIPCHECK(hipMemcpyAsync(&found, count_, sizeof found, hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  const uint32_t n = std::min(found, capacity);
  if (n && out) HIPCHECK(hipMemcpy(out, out_, n * sizeof(talus_mi_surface_contact), hipMemcpyDeviceToHost));
  return found;
}

}  // namespace tmi
