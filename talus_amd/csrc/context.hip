// context.hip -- context, device buffers, step orchestration and the C ABI
// of libtalus_mi.
//
// A context owns one HIP stream and every device buffer.  All bodies are
// concatenated into one set of SoA arrays and stepped by the same launches
// (one launch per kernel per substep pass, whatever the body count), each
// body with its own substep count, PCG scalars and convergence.
//
// Step flow per body follows deformable::step for a non-floating tet body
// (src/talus/deformable/solver/solver.cc:3472-5901): predictorInverts raises
// the substep count to 4 (3574-3583); per substep the alpha-point predictor
// (3680-3701), assembly + balance + PCG solve for the new acceleration, two
// re-assembled correction solves (ANCHORED_READS = 3, 3895-3950), velocity
// and position updates (4351-4368, 5829-5849) and commitPlastic (5851).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <functional>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "contact.h"
#include "factor.h"
#include "host.h"
#include "device.h"
#include "field.h"
#include "hexlattice.h"
#include "screen.h"
#include "kernels.h"
#include "tmath.h"

// Slot buffer of the allocation-free operator entry point: a file-scope
// device array, one copy per GPU for the life of the process.  Host code
// uses only the per-device address from hipGetSymbolAddress (the array's
// name is a host-side stand-in, not a device address).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(TALUS_MI_NOALLOC_SCRATCH_MIB) << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace tmi {

namespace {
thread_local std::string g_error;
}

int fail(int code, const std::string &message) {
  g_error = message;
  return code;
}

int prepare_body(Body &b) {
  const uint32_t lanes = uint32_t(b.materials.size());
  if (b.lame_set && (b.lambda.size() != b.tets || b.mu.size() != b.tets)) b.lame_set = false;
  b.lambda.resize(b.tets);
  b.mu.resize(b.tets);
  b.rho.assign(b.tets, 0.0);
  b.yield.resize(b.tets);
  b.hardening.resize(b.tets);
  b.plastic_max.resize(b.tets);
  b.tear.resize(b.tets);
  b.tau.resize(b.tets);
  b.flags.resize(b.tets);
  b.dminv.resize(9 * size_t(b.tets));
  b.volume.resize(b.tets);
  for (uint32_t e = 0; e < b.tets; ++e) {
    for (int c = 0; c < 4; ++c)
      if (b.tet[4 * size_t(e) + c] >= b.nodes) return fail(TALUS_MI_ERR_ARGUMENT, "tet node index out of range");
    if (!rest_frame(b.rest.data(), &b.tet[4 * size_t(e)], &b.dminv[9 * size_t(e)], &b.volume[e]))
      return fail(TALUS_MI_ERR_ARGUMENT, "tet " + std::to_string(e) + " is not positively oriented");
    const talus_mi_material &m = lane_material(b.materials.data(), lanes, b.lane[e]);
    if (!b.lame_set) precise_lame(m.youngs_modulus, m.poissons_ratio, &b.lambda[e], &b.mu[e]);
    if (b.density_mass) b.rho[e] = double(m.density);
    const bool plastic = std::isfinite(m.yield_stress) && m.yield_stress > 0.0f;
    b.yield[e] = plastic ? double(m.yield_stress) : INFINITY;
    b.hardening[e] = m.hardening_modulus > 0.0f ? double(m.hardening_modulus) : 0.0;
    b.plastic_max[e] = double(m.plastic_max);
    b.tear[e] = 1.0 + double(m.tear_strain);
    // stiffnessDamping (factor.cc:755-759): finite and positive, else 0
    b.tau[e] = std::isfinite(m.stiffness_damping) && m.stiffness_damping > 0.0f ? double(m.stiffness_damping) : 0.0;
    uint8_t f = 0;
    if (m.elastic_model == 1u) f |= FLAG_SNH;
    if (plastic) f |= FLAG_PLASTIC;
    if (std::isfinite(m.tear_strain)) f |= FLAG_TEAR;
    if (b.tau[e] > 0.0) f |= FLAG_DAMP;
    b.flags[e] = f;
  }
  if (b.density_mass) {
    b.mass.assign(b.nodes, 0.0);
    for (uint32_t e = 0; e < b.tets; ++e) {
      const double rho = double(lane_material(b.materials.data(), lanes, b.lane[e]).density);
      for (int c = 0; c < 4; ++c) b.mass[b.tet[4 * size_t(e) + c]] += rho * b.volume[e] * 0.25;
    }
  }
  return TALUS_MI_OK;
}

// ------------------------------------------------------------ plastic Newton

// One device operation of the plastic Newton: an evaluation (x_eval, the
// assembly with the return maps and the loading count, balance, |right|^2),
// a PCG solve, or an iterate update (NewtonVec).
struct NewtonOp {
  enum Kind : uint8_t { DONE = 0, EVAL = 1, SOLVE = 2, VEC = 3 };
  uint8_t kind = DONE;
  uint8_t first = 0;    // EVAL at the predictor (total 0) / SOLVE from the previous acceleration
  uint8_t tangent = 0;  // the consistent tangent bound (bindStiffness(consistent))
  uint8_t elastic = 0;  // EVAL without the return map
  int vec = 0;          // VEC: NewtonVec mode
  int key() const { return kind | first << 2 | tangent << 3 | elastic << 4 | vec << 5; }
};

// The plastic branch of deformable::step for one body and substep
// (solver.cc:3951-4250) as a resumable routine: next() follows the
// reference's control flow until it needs the device, returns that
// operation, and continues from there when called again with the
// operation's results filled in (squared, loading after an EVAL; moved,
// extent after a SOLVE).  The unknown, total, is the engine's acceleration
// after the first solve.
struct NewtonBody {
  static constexpr double MARGIN = 1.0e-3;     // LOADING_MARGIN (solver.cc:372)
  static constexpr double REDUCTION = 1.0e-2;  // PLASTIC_REDUCTION (374)
  static constexpr double RESOLUTION = 8.0;    // PLASTIC_RESOLUTION (380)
  static constexpr double FLOAT_EPSILON = 1.1920928955078125e-7;
  uint32_t cap = 40;    // plasticPasses
  double travel = 0.0;  // beta h^2: acceleration to position
  double squared = 0.0, moved = 0.0, extent = 0.0;
  uint32_t loading = 0;
  // StepReport counters
  uint32_t evaluations = 0, solves_all = 0, backtracks = 0, linear_reads = 0, handoffs = 0,
           fallbacks = 0, rejections = 0, capped = 0;

  NewtonOp next();

 private:
  static NewtonOp eval(bool first, bool tangent, bool elastic) {
    NewtonOp o;
    o.kind = NewtonOp::EVAL;
    o.first = first;
    o.tangent = tangent;
    o.elastic = elastic;
    return o;
  }
  static NewtonOp solve(bool first, bool tangent) {
    NewtonOp o;
    o.kind = NewtonOp::SOLVE;
    o.first = first;
    o.tangent = tangent;
    return o;
  }
  static NewtonOp vec(int mode) {
    NewtonOp o;
    o.kind = NewtonOp::VEC;
    o.vec = mode;
    return o;
  }
  int pc_ = 0;
  bool linear_allowed_ = true, tangent_ = false, predictor_unloaded_ = false, converged_ = false,
       handed_off_ = false, fell_back_ = false;
  double accepted_ = 0.0, opening_ = 0.0, lowest_ = 0.0, ranked_ = 0.0;
  uint32_t attempt_ = 0, solves_ = 0, loading_before_ = 0, read_ = 0;
};

// returns `op` and resumes after this point on the next call
#define TMI_AWAIT(op)    \
  do {                   \
    pc_ = __LINE__;      \
    return (op);         \
    case __LINE__:;      \
  } while (0)

NewtonOp NewtonBody::next() {
  constexpr double INF = std::numeric_limits<double>::infinity();
  switch (pc_) {
    case 0:
      for (;;) {
        accepted_ = INF;
        opening_ = 0.0;
        lowest_ = INF;
        loading_before_ = 0;
        solves_ = 0;
        predictor_unloaded_ = converged_ = handed_off_ = false;
        for (attempt_ = 0; attempt_ < cap; ++attempt_) {
          if (attempt_ < 2) tangent_ = attempt_ != 0;
          TMI_AWAIT(eval(attempt_ == 0, tangent_, false));
          ++evaluations;
          if (attempt_ != 0) {
            ranked_ = std::isfinite(squared) ? squared : INF;
            if (attempt_ == 1 || ranked_ < lowest_) {
              lowest_ = ranked_;
              TMI_AWAIT(vec(NV_SAVE));
            }
          }
          if (attempt_ == 0) predictor_unloaded_ = loading == 0;
          if (attempt_ == 1 && linear_allowed_ && predictor_unloaded_ && std::isfinite(squared) &&
              squared > accepted_) {
            // the elastic reads from the first solve's total
            ++linear_reads;
            tangent_ = false;
            for (read_ = 1; read_ < 3; ++read_) {
              TMI_AWAIT(eval(false, false, false));
              ++evaluations;
              if (loading != 0) {
                handed_off_ = true;
                break;
              }
              TMI_AWAIT(solve(false, false));
              ++solves_;
            }
            converged_ = !handed_off_;
            break;
          }
          if (attempt_ != 0 && !(squared <= accepted_)) {
            ++backtracks;
            TMI_AWAIT(vec(NV_BACKTRACK));
            continue;
          }
          if (attempt_ != 0 && loading_before_ == 0 && loading == 0) {
            converged_ = true;
            break;
          }
          if (attempt_ == 0) opening_ = squared;
          accepted_ = squared;
          loading_before_ = loading;
          TMI_AWAIT(solve(attempt_ == 0, tangent_));
          ++solves_;
          if (attempt_ == 0) TMI_AWAIT(vec(NV_DX_FROM_A));
          if (solves_ > 1 && travel * moved <= RESOLUTION * FLOAT_EPSILON * extent) {
            converged_ = true;
            break;
          }
        }
        solves_all += solves_;
        if (handed_off_) {
          ++handoffs;
          linear_allowed_ = false;
          TMI_AWAIT(vec(NV_RESET));
          continue;
        }
        if (!converged_) {
          // the elastic scheme from total 0, accepted when it loads nothing
          // and reduces the residual 100-fold
          tangent_ = false;
          TMI_AWAIT(vec(NV_RESET));
          fell_back_ = true;
          for (read_ = 0; read_ < 3; ++read_) {
            TMI_AWAIT(eval(read_ == 0, false, read_ == 0));
            ++evaluations;
            if (loading != 0) {
              fell_back_ = false;
              break;
            }
            TMI_AWAIT(solve(read_ == 0, false));
            ++solves_all;
          }
          if (fell_back_) {
            TMI_AWAIT(eval(false, false, false));
            ++evaluations;
            fell_back_ = loading == 0 && squared <= REDUCTION * REDUCTION * opening_;
          }
          if (fell_back_) {
            ++fallbacks;
            converged_ = true;
          } else {
            // keep the iterate with the lowest residual
            ++rejections;
            tangent_ = true;
            TMI_AWAIT(vec(NV_RESTORE));
            TMI_AWAIT(eval(false, true, false));
            ++evaluations;
          }
        }
        capped += converged_ ? 0u : 1u;
        break;
      }
      pc_ = -1;
      return NewtonOp{};
    default:
      return NewtonOp{};
  }
}

#undef TMI_AWAIT

// ------------------------------------------------------------------ engine

class EngineBase {
 public:
  virtual ~EngineBase() = default;
  virtual void build(const std::vector<Body> &bodies) = 0;
  virtual void download(std::vector<Body> &bodies) = 0;
  virtual void step(const talus_mi_step_params &sp, talus_mi_report &rep,
                    std::vector<talus_mi_event> &events, uint32_t step_index,
                    std::vector<uint8_t> &torn) = 0;
  // the last step's report of body b alone (zero before the first step)
  virtual void body_report(uint32_t b, talus_mi_report &out) const = 0;
  // device views of body b's positions and velocities (FP64, 3 per node)
  // and the stream whose work writes them
  virtual void device_state(uint32_t b, const double **x, const double **v, hipStream_t *s) const = 0;
  virtual void state_set(uint32_t b, const double *x, const double *v, const double *a) = 0;
  virtual void state_get(uint32_t b, double *x, double *v, double *a) = 0;
  virtual void rotations_get(uint32_t b, double *R) = 0;
  virtual void rotations_set(uint32_t b, const double *R) = 0;
  virtual void frames_get(uint32_t b, double *W) = 0;
  virtual void frames_set(uint32_t b, const double *W) = 0;
  virtual void plastic_get(uint32_t b, double *p) = 0;
  virtual void plastic_set(uint32_t b, const double *p) = 0;
  virtual void eval_forces(uint32_t b, const double *x, double *tf, double *te, double *nf,
                           double *R) = 0;
  virtual void apply_system(uint32_t b, double inertia, double step, const double *dv,
                            double *y) = 0;
  // the operator on device vectors with a caller-given slot buffer; launches
  // only (no allocation, no host copy, no wait) on stream s
  virtual void apply_system_device(uint32_t b, double inertia, double step, const double *dv,
                                   double *y, void *slots, hipStream_t s) = 0;
  virtual void apply_system_stressed(uint32_t b, const double *x, double inertia, double step,
                                     const double *dv, double *y) = 0;
  virtual void apply_system_damped(uint32_t b, const double *x, double inertia, double step, double damp,
                                   const double *dv, double *y) = 0;
  virtual void time_kernel(uint32_t kind, uint32_t repeats, talus_mi_kernel_timing &out) = 0;
  virtual void factor_solve(uint32_t b, double inertia, double step, const double *r,
                            double *z) = 0;
  virtual void factor_info(talus_mi_factor_info &info) = 0;
  virtual void factor_profile(std::vector<talus_mi_factor_kernel> &out) = 0;
  virtual void body_frame(uint32_t b, double *frame) = 0;
  virtual void tune(uint32_t repeats, talus_mi_tune_report &report) = 0;
  virtual void state_all_get(std::vector<double> &xva) = 0;
  virtual void state_all_set(const std::vector<double> &xva) = 0;
  virtual void rollout_backward(const talus_mi_step_params &sp,
                                const std::vector<std::vector<double>> &ckpt,
                                const std::vector<double> &final_state, uint32_t steps,
                                uint32_t every, uint32_t body, const double *x_bar,
                                const double *v_bar, double tolerance, double *lambda_bar,
                                double *mu_bar, double *rest_bar, double *x0_bar, double *v0_bar,
                                double *a0_bar) = 0;
  virtual std::vector<uint8_t> plan() = 0;
  // supernode leaf size of the next analysis (nodes per leaf front)
  virtual void set_leaf(uint32_t leaf) = 0;
  virtual bool set_plan(const std::vector<uint8_t> &plan) = 0;
};

template <class Real>
class Engine final : public EngineBase {
 public:
  explicit Engine(hipStream_t s) : stream_(s) {}
  ~Engine() override { release(); }

  void build(const std::vector<Body> &bodies) override;
  void download(std::vector<Body> &bodies) override;
  void step(const talus_mi_step_params &sp, talus_mi_report &rep,
            std::vector<talus_mi_event> &events, uint32_t step_index,
            std::vector<uint8_t> &torn) override;
  void state_set(uint32_t b, const double *x, const double *v, const double *a) override;
  void state_get(uint32_t b, double *x, double *v, double *a) override;
  void rotations_get(uint32_t b, double *R) override { planes_get(d_.R, b, R); }
  void rotations_set(uint32_t b, const double *R) override { planes_set(d_.R, b, R); }
  void frames_get(uint32_t b, double *W) override { planes_get(d_.W, b, W); }
  void frames_set(uint32_t b, const double *W) override { planes_set(d_.W, b, W); }
  void planes_get(const Real *base, uint32_t b, double *out);
  void planes_set(Real *base, uint32_t b, const double *in);
  void plastic_get(uint32_t b, double *p) override;
  void plastic_set(uint32_t b, const double *p) override;
  void eval_forces(uint32_t b, const double *x, double *tf, double *te, double *nf,
                   double *R) override;
  void apply_system(uint32_t b, double inertia, double step, const double *dv,
                    double *y) override;
  void apply_system_device(uint32_t b, double inertia, double step, const double *dv, double *y,
                           void *slots, hipStream_t s) override {
    Launch<Real>::apply_body(d_, t0_[b], t0_[b] + tt_[b], n0_[b], n0_[b] + nn_[b], inc_start_[n0_[b]], dv, y,
                             inertia, step * step, static_cast<Real *>(slots), s);
  }
  void apply_system_stressed(uint32_t b, const double *x, double inertia, double step, const double *dv,
                             double *y) override;
  void apply_system_damped(uint32_t b, const double *x, double inertia, double step, double damp,
                           const double *dv, double *y) override;
  void body_report(uint32_t b, talus_mi_report &out) const override {
    if (b < body_rep_.size())
      out = body_rep_[b];
    else
      std::memset(&out, 0, sizeof out);
  }
  void device_state(uint32_t b, const double **x, const double **v, hipStream_t *s) const override {
    *x = d_.x + 3 * size_t(n0_[b]);
    *v = d_.v + 3 * size_t(n0_[b]);
    *s = stream_;
  }
  void time_kernel(uint32_t kind, uint32_t repeats, talus_mi_kernel_timing &out) override;
  void factor_solve(uint32_t b, double inertia, double step, const double *r, double *z) override;
  void factor_info(talus_mi_factor_info &info) override;
  void factor_profile(std::vector<talus_mi_factor_kernel> &out) override;
  void body_frame(uint32_t b, double *frame) override;
  void tune(uint32_t repeats, talus_mi_tune_report &report) override;
  void state_all_get(std::vector<double> &xva) override {
    const size_t n3 = 3 * size_t(d_.N);
    xva.resize(3 * n3);
    HIPCHECK(hipStreamSynchronize(stream_));
    HIPCHECK(hipMemcpy(xva.data(), d_.x, n3 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(xva.data() + n3, d_.v, n3 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(xva.data() + 2 * n3, d_.a, n3 * sizeof(double), hipMemcpyDeviceToHost));
  }
  void state_all_set(const std::vector<double> &xva) override {
    const size_t n3 = 3 * size_t(d_.N);
    HIPCHECK(hipStreamSynchronize(stream_));
    HIPCHECK(hipMemcpy(d_.x, xva.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_.v, xva.data() + n3, n3 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_.a, xva.data() + 2 * n3, n3 * sizeof(double), hipMemcpyHostToDevice));
  }
  void rollout_backward(const talus_mi_step_params &sp, const std::vector<std::vector<double>> &ckpt,
                        const std::vector<double> &final_state, uint32_t steps, uint32_t every,
                        uint32_t body, const double *x_bar, const double *v_bar, double tolerance,
                        double *lambda_bar, double *mu_bar, double *rest_bar, double *x0_bar,
                        double *v0_bar, double *a0_bar) override;
  void set_leaf(uint32_t leaf) override { leaf_ = leaf; }
  std::vector<uint8_t> plan() override {
    if (!sym_ready_) analyze_factor();
    return plan_;
  }
  bool set_plan(const std::vector<uint8_t> &plan) override {
    if (!sym_ready_) analyze_factor();
    if (plan.size() != plan_.size()) return false;
    for (uint8_t p : plan)
      if (p > 2) return false;
    plan_ = plan;
    ++graph_epoch_;
    return true;
  }

 private:
  template <class T>
  T *dev(size_t n) {
    T *p = nullptr;
    HIPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    owned_.push_back(p);
    return p;
  }
  template <class T>
  T *dev_copy(const std::vector<T> &h) {
    T *p = dev<T>(h.size());
    if (!h.empty()) HIPCHECK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  template <class T>
  T *pinned(size_t n) {
    T *p = nullptr;
    HIPCHECK(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault));
    pinned_.push_back(p);
    return p;
  }
  void release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto &a : graph_)
      for (auto &b : a)
        for (Graph &g : b) {
          if (g.exec) (void)hipGraphExecDestroy(g.exec);
          g = Graph{};
        }
    ++graph_epoch_;
    for (void *p : owned_) (void)hipFree(p);
    for (void *p : pinned_) (void)hipHostFree(p);
    owned_.clear();
    pinned_.clear();
    built_ = false;
  }
  void upload_params(const std::vector<StepParams> &params, size_t slot);
  // rhs: the right-hand side (FP64, 3N) in place of d_.right (refinement's
  // correction solves)
  void solve(bool first, uint32_t limit, uint32_t solve_index, bool tangent = false,
             double *rhs = nullptr);
  void all_active();
  // block Cholesky preconditioner
  void analyze_factor();
  // Partial refreshes launch only the refreshed bodies' factor tasks: every
  // task list holds each launch's tasks body by body (fronts are numbered
  // body-major), so a body's tasks in a launch are one contiguous range.
  // Compact copies of a partial refresh's lists are staged in pinned memory
  // and uploaded once per refresh.
  struct TaskSpan {
    const void *dev;  // first task on the device
    uint32_t count;
  };
  bool compactable_ = false;
  uint64_t *hcompact64_ = nullptr, *dcompact64_ = nullptr;
  uint32_t *hcompact32_ = nullptr, *dcompact32_ = nullptr;
  size_t n64_ = 0, n32_ = 0;
  const std::vector<int> *compact_mask_ = nullptr;
  template <class T>
  TaskSpan compact(const std::vector<T> &tasks, const T *dev, uint32_t s0, uint32_t s1);
  // per-launch HIP events of one factorization (factor_profile)
  struct FactorProfile {
    struct Launch {
      uint32_t kind;
      uint64_t workgroups;
      double mfma, useful;
      hipEvent_t a, b;
    };
    std::vector<Launch> launches;
  };
  void factorize(const std::vector<double> &reach, const std::vector<double> &inertia, const std::vector<double> *damp,
                 const std::vector<int> *mask = nullptr, FactorProfile *prof = nullptr);
  void precondition(int parity);
  void forward_level(uint32_t l, int mode, int parity, const Real *rin = nullptr);
  void backward_level(uint32_t l, int mode, int parity);
  template <class T>
  T *upload(const std::vector<T> &h) { return dev_copy(h); }

  hipStream_t stream_;
  std::vector<void *> owned_, pinned_;
  bool built_ = false;
  Dev<Real> d_{};
  std::vector<uint32_t> n0_, t0_, nn_, tt_, nt0_, nt1_, tt0_, tt1_;
  // pinned staging
  StepParams *hparams_ = nullptr;
  size_t hparams_slots_ = 0;
  int *hflags_ = nullptr;
  uint32_t *hiters_ = nullptr;
  double *hrr_ = nullptr, *htarget_ = nullptr;
  size_t hsolve_slots_ = 0;
  double *hpartial_ = nullptr, *hepartial_ = nullptr;
  uint32_t *hyields_ = nullptr;
  Real *hvec_ = nullptr;  // 3N + 12T staging of Real
  double *scratch64_ = nullptr;
  bool any_plastic_ = false, any_tear_ = false, any_damped_ = false;
  uint64_t n_snh_ = 0, n_plastic_ = 0, n_flowing_ = 0;  // flowing: plastic neo-Hookean
  uint64_t tn_total_ = 0;  // entries of the tet tiles' node lists
  uint32_t applications_ = 0;  // preconditioner applications this step
  // host copies of the incidence CSR: node ranges and the CSR position of
  // every (tet, corner) slot code 4e+c
  std::vector<uint32_t> inc_start_, slot_pos_;
  double *hnorms_ = nullptr, *d_norms_ = nullptr;  // corrections' convergence
  // FP64 iterative refinement: pinned copies of the per-body test and
  // relative residual, and the step's tallies
  int *hrdone_ = nullptr;
  double *hrrel_ = nullptr;
  struct RefineTally {
    uint32_t corrections = 0, unmet = 0;
    double worst = 0.0;
    std::vector<uint32_t> body_corrections, body_unmet;
    std::vector<double> body_worst;
  };
  // refine the solve just made; increment: it solved for an increment added
  // to a (d_.rbase keeps that increment's FP64 sum)
  void refine(const talus_mi_step_params &sp, uint32_t limit, uint32_t pass, const std::vector<uint8_t> &mask,
              bool increment, std::vector<uint32_t> &solve_pass, std::vector<std::vector<uint8_t>> &solve_active,
              uint32_t &solve_index, RefineTally &tally);
  void ensure_refine();
  // plastic Newton: bodies with a yielding corotated tet, pinned staging of
  // its per-tile norms, loading counts and per-operation parameters
  bool frames_ = false;  // the factor follows each body's rotation (bodyRotation)
  std::vector<uint8_t> newton_body_;
  double *hnnorm_ = nullptr;
  uint32_t *hloading_ = nullptr;
  static constexpr size_t NEWTON_SLOTS = 32;
  StepParams *hnparams_ = nullptr;
  size_t nslot_ = 0;
  void upload_newton(const std::vector<StepParams> &params);
  void ensure_solve_slot(uint32_t solve_index);
  void newton_pass(const std::vector<uint32_t> &bodies, const std::vector<StepParams> &params,
                   uint32_t cap, uint32_t limit, bool diag, uint32_t pass,
                   std::vector<uint32_t> &solve_pass, std::vector<std::vector<uint8_t>> &solve_active,
                   uint32_t &solve_index, talus_mi_report &rep);
  struct Recorder {
    std::vector<double *> buffers;  // per pass: x_n, v_n, a_n, a' (3N each)
    std::vector<std::vector<StepParams>> params;
  };
  Recorder *recorder_ = nullptr;
  // captured PCG iterations, [parity][factor][tangent]; rebuilt when the
  // engine, the factor's analysis or the solve plan change (graph_epoch_)
  struct Graph {
    hipGraphExec_t exec = nullptr;
    uint64_t epoch = ~0ull;
  };
  Graph graph_[2][2][2];
  uint64_t graph_epoch_ = 0;
  bool graphs_ = true;

 public:
  void use_graphs(bool on) { graphs_ = on; }

 private:
  // factor state
  const std::vector<Body> *bodies_ = nullptr;
  bool sym_ready_ = false, fac_ready_ = false, fac_ok_ = false, use_factor_ = false;
  std::vector<double> fac_reach_, fac_inertia_, fac_damp_;
  FactorTables tab_;
  uint32_t leaf_ = DEFAULT_LEAF;  // supernode leaf size (nodes)
  std::vector<uint8_t> plan_;  // per level: forward, backward kernel shape
  FactorDev fd_{};
  uint64_t *d_pair_target_ = nullptr;
  uint32_t *d_pair_ld_ = nullptr, *d_pair_node_ = nullptr, *d_pair_start_ = nullptr,
           *d_contrib_ = nullptr, *d_pair_body_ = nullptr, *d_parent_ = nullptr,
           *d_lvl_fronts_ = nullptr, *d_potrf_ = nullptr, *d_node_body_ = nullptr;
  uint64_t *d_ea_ = nullptr, *d_trsm_ = nullptr, *d_syrk_ = nullptr, *d_trtri_ = nullptr,
           *d_w21_ = nullptr, *d_pack_ = nullptr, *d_fwd_ = nullptr, *d_bwd_ = nullptr,
           *d_fwl_ = nullptr, *d_bwl_ = nullptr, *d_fwr_ = nullptr, *d_bwr_ = nullptr,
           *d_fws_ = nullptr, *d_bws_ = nullptr, *d_zero_ = nullptr, *d_gat_all_ = nullptr,
           *d_gat_big_ = nullptr, *d_xg_all_ = nullptr, *d_xg_big_ = nullptr;
  using GT = typename GStore<Real>::type;  // storage of the solve matrix G
  GT *d_gr_ = nullptr, *d_gc_ = nullptr;
  double *d_body_reach_ = nullptr, *d_body_inertia_ = nullptr, *d_body_damp_ = nullptr;
  int *hfailed_ = nullptr;
  int *d_refresh_ = nullptr;
  // early refresh of single bodies' factors: a body whose solves needed at
  // least REFRESH_ITERS iterations in a step is refactored at the next step
  // (at most every REFRESH_COOLDOWN steps); a capped solve refreshes its body
  static constexpr uint32_t REFRESH_ITERS = 4, REFRESH_COOLDOWN = 2;
  std::vector<int> want_refresh_;
  std::vector<talus_mi_report> body_rep_;  // the last step's report of each body
  std::vector<uint64_t> next_refresh_;
  uint64_t steps_taken_ = 0;
  uint32_t refactors_ = 0;
  double factor_seconds_ = 0.0;

 public:
  const FactorTables &tables() const { return tab_; }
  double last_factor_seconds() const { return factor_seconds_; }
};

template <class Real>
void Engine<Real>::build(const std::vector<Body> &bodies) {
  release();
  bodies_ = &bodies;
  sym_ready_ = fac_ready_ = fac_ok_ = false;
  hnorms_ = d_norms_ = nullptr;
  hrdone_ = nullptr;
  hrrel_ = nullptr;
  want_refresh_.assign(bodies.size(), 0);
  next_refresh_.assign(bodies.size(), 0);
  fd_ = FactorDev{};
  const uint32_t B = uint32_t(bodies.size());
  uint64_t N = 0, T = 0;
  n0_.assign(B, 0);
  t0_.assign(B, 0);
  nn_.assign(B, 0);
  tt_.assign(B, 0);
  for (uint32_t b = 0; b < B; ++b) {
    n0_[b] = uint32_t(N);
    t0_[b] = uint32_t(T);
    nn_[b] = bodies[b].nodes;
    tt_[b] = bodies[b].tets;
    N += bodies[b].nodes;
    T += bodies[b].tets;
  }
  if (T >= (1ull << 30) || N >= (1ull << 31)) throw std::length_error("too many tets or nodes");
  d_ = Dev<Real>{};
  d_.N = uint32_t(N);
  d_.T = uint32_t(T);
  d_.B = B;
  // tiles
  std::vector<uint32_t> ntb, ntbeg, ntend, ttb, ttbeg, ttend;
  nt0_.assign(B, 0);
  nt1_.assign(B, 0);
  tt0_.assign(B, 0);
  tt1_.assign(B, 0);
  for (uint32_t b = 0; b < B; ++b) {
    nt0_[b] = uint32_t(ntb.size());
    for (uint32_t s = 0; s < nn_[b]; s += TILE) {
      ntb.push_back(b);
      ntbeg.push_back(n0_[b] + s);
      ntend.push_back(n0_[b] + std::min<uint32_t>(s + TILE, nn_[b]));
    }
    nt1_[b] = uint32_t(ntb.size());
    tt0_[b] = uint32_t(ttb.size());
    for (uint32_t s = 0; s < tt_[b]; s += TILE) {
      ttb.push_back(b);
      ttbeg.push_back(t0_[b] + s);
      ttend.push_back(t0_[b] + std::min<uint32_t>(s + TILE, tt_[b]));
    }
    tt1_[b] = uint32_t(ttb.size());
  }
  d_.n_ntiles = uint32_t(ntb.size());
  d_.n_ttiles = uint32_t(ttb.size());
  d_.nt_body = dev_copy(ntb);
  d_.nt_begin = dev_copy(ntbeg);
  d_.nt_end = dev_copy(ntend);
  d_.tt_body = dev_copy(ttb);
  d_.tt_begin = dev_copy(ttbeg);
  d_.tt_end = dev_copy(ttend);
  d_.body_nt0 = dev_copy(nt0_);
  d_.body_nt1 = dev_copy(nt1_);
  // nodes
  std::vector<double> X(3 * N), mass(N), x(3 * N), v(3 * N), a(3 * N);
  std::vector<uint8_t> anchored(N);
  std::vector<uint32_t> inc_start(N + 1, 0);
  slot_pos_.assign(4 * T, 0);
  std::vector<uint4> conn(T);
  std::vector<double> dminv(9 * T), vol64(T), lam64(T), mu64(T), rho(T), yield(T), hard(T), pmax(T),
      tear(T), plastic(10 * T);
  std::vector<Real> vol(T), lam(T), mu(T), R(9 * T), W(9 * T), tau(T);
  std::vector<uint8_t> flags(T);
  any_plastic_ = any_tear_ = any_damped_ = false;
  n_snh_ = n_plastic_ = n_flowing_ = 0;
  newton_body_.assign(B, 0);
  for (uint32_t b = 0; b < B; ++b) {
    const Body &bd = bodies[b];
    const size_t n0 = n0_[b], t0 = t0_[b];
    std::copy(bd.rest.begin(), bd.rest.end(), X.begin() + 3 * n0);
    std::copy(bd.mass.begin(), bd.mass.end(), mass.begin() + n0);
    std::copy(bd.anchored.begin(), bd.anchored.end(), anchored.begin() + n0);
    std::copy(bd.x.begin(), bd.x.end(), x.begin() + 3 * n0);
    std::copy(bd.v.begin(), bd.v.end(), v.begin() + 3 * n0);
    std::copy(bd.a.begin(), bd.a.end(), a.begin() + 3 * n0);
    for (uint32_t e = 0; e < bd.tets; ++e) {
      const size_t g = t0 + e;
      const uint32_t *tn = &bd.tet[4 * size_t(e)];
      conn[g] = make_uint4(uint32_t(n0 + tn[0]), uint32_t(n0 + tn[1]), uint32_t(n0 + tn[2]),
                           uint32_t(n0 + tn[3]));
      for (int c = 0; c < 4; ++c) ++inc_start[n0 + tn[c] + 1];
      for (int k = 0; k < 9; ++k) {
        dminv[size_t(k) * T + g] = bd.dminv[9 * size_t(e) + k];
        R[size_t(k) * T + g] = Real(bd.R[9 * size_t(e) + k]);
        W[size_t(k) * T + g] = Real(bd.W[9 * size_t(e) + k]);
      }
      for (int k = 0; k < 10; ++k) plastic[size_t(k) * T + g] = bd.plastic[10 * size_t(e) + k];
      vol64[g] = bd.volume[e];
      rho[g] = bd.rho[e];
      lam64[g] = bd.lambda[e];
      mu64[g] = bd.mu[e];
      vol[g] = Real(bd.volume[e]);
      lam[g] = Real(bd.lambda[e]);
      mu[g] = Real(bd.mu[e]);
      yield[g] = bd.yield[e];
      hard[g] = bd.hardening[e];
      pmax[g] = bd.plastic_max[e];
      tear[g] = bd.tear[e];
      tau[g] = Real(bd.tau[e]);
      flags[g] = bd.flags[e];
      any_damped_ = any_damped_ || (bd.flags[e] & FLAG_DAMP);
      any_plastic_ = any_plastic_ || (bd.flags[e] & FLAG_PLASTIC);
      if ((bd.flags[e] & FLAG_PLASTIC) && !(bd.flags[e] & FLAG_SNH)) newton_body_[b] = 1;
      n_snh_ += (bd.flags[e] & FLAG_SNH) ? 1u : 0u;
      n_plastic_ += (bd.flags[e] & FLAG_PLASTIC) ? 1u : 0u;
      n_flowing_ += ((bd.flags[e] & FLAG_SNH) && (bd.flags[e] & FLAG_PLASTIC)) ? 1u : 0u;
      any_tear_ = any_tear_ || (bd.flags[e] & FLAG_TEAR);
    }
  }
  for (size_t n = 0; n < N; ++n) inc_start[n + 1] += inc_start[n];
  {
    // slot code 4e+c -> its CSR position; codes ascend within each node
    std::vector<uint32_t> cursor(inc_start.begin(), inc_start.end() - 1);
    const uint32_t *cn = reinterpret_cast<const uint32_t *>(conn.data());
    for (size_t s = 0; s < 4 * T; ++s) slot_pos_[s] = cursor[cn[s]]++;
  }
  d_.X = dev_copy(X);
  d_.mass = dev_copy(mass);
  d_.anchored = dev_copy(anchored);
  d_.inc_start = dev_copy(inc_start);
  d_.slot_pos = reinterpret_cast<const uint4 *>(dev_copy(slot_pos_));
  inc_start_.assign(inc_start.begin(), inc_start.end());
  d_.x = dev_copy(x);
  d_.v = dev_copy(v);
  d_.a = dev_copy(a);
  d_.carried = dev<double>(3 * N);
  d_.shifted = dev<double>(3 * N);
  d_.weighted = dev<double>(3 * N);
  d_.right = dev<double>(3 * N);
  d_.u_eval = dev<double>(3 * N);
  d_.P = dev<Real>(6 * N);
  d_.x0 = dev<Real>(3 * N);
  d_.dx = dev<Real>(3 * N);
  d_.r = dev<Real>(3 * N);
  d_.z = dev<Real>(3 * N);
  d_.p = dev<Real>(3 * N);
  d_.q = dev<Real>(3 * N);
  d_.conn = dev_copy(conn);
  {
    // each tet tile's nodes, once each in ascending order, and the tets'
    // corners as positions in that list
    std::vector<uint32_t> tn_start(1, 0), tn_node, scratch;
    std::vector<uint2> conn_l(T);
    uint32_t tn_max = 0;
    for (size_t t = 0; t < ttb.size(); ++t) {
      scratch.clear();
      for (uint32_t e = ttbeg[t]; e < ttend[t]; ++e) {
        const uint4 c = conn[e];
        scratch.insert(scratch.end(), {c.x, c.y, c.z, c.w});
      }
      std::sort(scratch.begin(), scratch.end());
      scratch.erase(std::unique(scratch.begin(), scratch.end()), scratch.end());
      auto at = [&](uint32_t n) {
        return uint32_t(std::lower_bound(scratch.begin(), scratch.end(), n) - scratch.begin());
      };
      for (uint32_t e = ttbeg[t]; e < ttend[t]; ++e) {
        const uint4 c = conn[e];
        conn_l[e] = make_uint2(at(c.x) | (at(c.y) << 16), at(c.z) | (at(c.w) << 16));
      }
      tn_node.insert(tn_node.end(), scratch.begin(), scratch.end());
      tn_start.push_back(uint32_t(tn_node.size()));
      tn_max = std::max(tn_max, uint32_t(scratch.size()));
    }
    static_assert(4 * TILE <= 65536, "tile node positions fit 16 bits");
    d_.tn_start = dev_copy(tn_start);
    d_.tn_node = dev_copy(tn_node);
    d_.conn_l = dev_copy(conn_l);
    d_.tn_max = tn_max;
    tn_total_ = tn_node.size();
  }
  d_.dminv = dev_copy(dminv);
  if constexpr (sizeof(Real) == sizeof(double)) {
    d_.dminv_r = reinterpret_cast<const Real *>(d_.dminv);
  } else {
    std::vector<Real> dminv_r(dminv.begin(), dminv.end());
    d_.dminv_r = dev_copy(dminv_r);
  }
  d_.vol = dev_copy(vol);
  d_.lam = dev_copy(lam);
  d_.mu = dev_copy(mu);
  d_.vol64 = dev_copy(vol64);
  d_.rho = dev_copy(rho);
  d_.lam64 = dev_copy(lam64);
  d_.mu64 = dev_copy(mu64);
  d_.flags = dev_copy(flags);
  d_.yield = dev_copy(yield);
  d_.hard = dev_copy(hard);
  d_.pmax = dev_copy(pmax);
  d_.tear = dev_copy(tear);
  d_.R = dev_copy(R);
  d_.W = dev_copy(W);
  d_.plastic = dev_copy(plastic);
  if (any_damped_) {
    d_.tau = dev_copy(tau);
    d_.fit = dev<Real>(size_t(FIT_PLANES) * T);
    HIPCHECK(hipMemset(d_.fit, 0, size_t(FIT_PLANES) * T * sizeof(Real)));
  }
  {
    std::vector<Real> lin(8 * size_t(T), Real(0));
    std::fill(lin.begin(), lin.begin() + T, Real(1));
    d_.lin = dev_copy(lin);
  }
  d_.loading = dev<uint32_t>(B);
  d_.fref = dev_copy(x);
  {
    std::vector<double> eye(9 * size_t(B), 0.0);
    for (uint32_t b = 0; b < B; ++b) eye[9 * b] = eye[9 * b + 4] = eye[9 * b + 8] = 1.0;
    d_.frame = dev_copy(eye);
  }
  d_.fpart = dev<double>(9 * size_t(d_.n_ntiles));
  d_.fcent = dev<double>(6 * size_t(B));
  d_.rf = dev<Real>(3 * N);
  d_.lowest = dev<double>(3 * N);
  d_.nnorm = dev<double>(3 * size_t(d_.n_ntiles));
  d_.slot_f = dev<Real>(12 * T);
  d_.slot_d = dev<Real>(24 * T);
  d_.energy = dev<Real>(T);
  d_.yflag = dev<uint8_t>(T);
  d_.tflag = dev<uint8_t>(T);
  d_.partial = dev<double>(size_t(d_.n_ntiles) * NPART);
  d_.epartial = dev<double>(d_.n_ttiles);
  d_.sq = dev<double>(2 * size_t(B));
  d_.done = dev<int>(2 * size_t(B));
  d_.iters = dev<uint32_t>(B);
  d_.target2 = dev<double>(B);
  d_.rr_last = dev<double>(B);
  d_.inverts = dev<int>(B);
  d_.yields = dev<uint32_t>(B);
  d_.params = dev<StepParams>(B);
  HIPCHECK(hipMemset(d_.yflag, 0, std::max<size_t>(T, 1)));
  HIPCHECK(hipMemset(d_.tflag, 0, std::max<size_t>(T, 1)));
  HIPCHECK(hipMemset(d_.epartial, 0, std::max<size_t>(d_.n_ttiles, 1) * sizeof(double)));
  scratch64_ = dev<double>(9 * N);  // three node vectors (eval_forces, apply_system)
  hparams_slots_ = 8;
  hparams_ = pinned<StepParams>(hparams_slots_ * B);
  hflags_ = pinned<int>(B);
  hsolve_slots_ = 3 * 8;
  hiters_ = pinned<uint32_t>(hsolve_slots_ * B);
  hrr_ = pinned<double>(hsolve_slots_ * B);
  htarget_ = pinned<double>(hsolve_slots_ * B);
  hpartial_ = pinned<double>(size_t(d_.n_ntiles) * NPART);
  hepartial_ = pinned<double>(d_.n_ttiles);
  hyields_ = pinned<uint32_t>(B);
  hvec_ = pinned<Real>(3 * N + 12 * T + 9 * T);
  hnnorm_ = pinned<double>(3 * size_t(d_.n_ntiles));
  hloading_ = pinned<uint32_t>(B);
  hnparams_ = pinned<StepParams>(NEWTON_SLOTS * B);
  nslot_ = 0;
  {
    std::vector<uint32_t> nb(N);
    for (uint32_t b = 0; b < B; ++b)
      for (uint32_t n = 0; n < nn_[b]; ++n) nb[n0_[b] + n] = b;
    d_node_body_ = dev_copy(nb);
  }
  hfailed_ = pinned<int>(B);
  HIPCHECK(hipDeviceSynchronize());
  built_ = true;
}

// Symbolic analysis of every body and upload of the factor tables
// (analyzeSystem, factor.cc:549-565).
template <class Real>
void Engine<Real>::analyze_factor() {
  const std::vector<Body> &bodies = *bodies_;
  std::vector<Symbolic> sym(bodies.size());
  for (size_t b = 0; b < bodies.size(); ++b) sym[b] = analyze(bodies[b], leaf_);
  build_tables(sym, bodies, n0_, t0_, tab_);
  const FactorTables &t = tab_;
  fd_.fronts = t.fronts;
  fd_.k = upload(t.k);
  fd_.m = upload(t.m);
  fd_.body = upload(t.body);
  fd_.rows_off = upload(t.rows_off);
  fd_.map_off = upload(t.map_off);
  fd_.child_start = upload(t.child_start);
  fd_.children = upload(t.children);
  fd_.rows = upload(t.rows);
  fd_.map = upload(t.map);
  fd_.off = upload(t.off);
  fd_.minv_off = upload(t.minv_off);
  fd_.vec_off = upload(t.vec_off);
  fd_.yoff = upload(t.yoff);
  fd_.g_off = upload(t.g_off);
  fd_.src_start = upload(t.src_start);
  fd_.src = upload(t.src);
  fd_.pool = dev<double>(t.pool);
  fd_.minv = dev<double>(t.minv_pool);
  fd_.vec = dev<double>(t.vec_pool);
  fd_.y = dev<double>(t.y_pool);
  fd_.gscale = dev<double>(t.y_pool);
  fd_.xsol = dev<double>(3 * size_t(d_.N));
  fd_.failed = dev<int>(d_.B);
  d_refresh_ = dev<int>(d_.B);
  fd_.refresh = d_refresh_;
  {
    // tasks of one launch are body-major in every list (partial refreshes
    // rely on it); staging for compact lists
    auto major = [&](const auto &tasks, const std::vector<uint32_t> &start) {
      for (size_t r = 0; r + 1 < start.size(); ++r)
        for (uint32_t q = start[r] + 1; q < start[r + 1]; ++q)
          if (t.body[uint32_t(tasks[q])] < t.body[uint32_t(tasks[q - 1])]) return false;
      return true;
    };
    const std::vector<uint32_t> one64 = {0u, uint32_t(t.w21_tasks.size())},
                                one_pack = {0u, uint32_t(t.pack_tasks.size())},
                                one_zero = {0u, uint32_t(t.zero_tasks.size())};
    compactable_ = major(t.zero_tasks, one_zero) && major(t.ea_tasks, t.ea_start) &&
                   major(t.potrf_tasks, t.potrf_start) && major(t.trsm_tasks, t.trsm_start) &&
                   major(t.syrk_tasks, t.syrk_start) && major(t.trtri_tasks, t.trtri_start) &&
                   major(t.w21_tasks, one64) && major(t.pack_tasks, one_pack);
    const size_t n64 = t.zero_tasks.size() + t.ea_tasks.size() + t.trsm_tasks.size() + t.syrk_tasks.size() +
                       t.trtri_tasks.size() + t.w21_tasks.size() + t.pack_tasks.size();
    hcompact64_ = pinned<uint64_t>(n64);
    dcompact64_ = dev<uint64_t>(n64);
    hcompact32_ = pinned<uint32_t>(t.potrf_tasks.size());
    dcompact32_ = dev<uint32_t>(t.potrf_tasks.size());
  }
  d_zero_ = upload(t.zero_tasks);
  d_parent_ = upload(t.parent);
  d_pair_target_ = upload(t.pair_target);
  d_pair_ld_ = upload(t.pair_ld);
  d_pair_node_ = upload(t.pair_node);
  d_pair_body_ = upload(t.pair_body);
  d_pair_start_ = upload(t.pair_start);
  d_contrib_ = upload(t.contrib);
  d_lvl_fronts_ = upload(t.lvl_fronts);
  d_ea_ = upload(t.ea_tasks);
  d_potrf_ = upload(t.potrf_tasks);
  d_trsm_ = upload(t.trsm_tasks);
  d_syrk_ = upload(t.syrk_tasks);
  d_trtri_ = upload(t.trtri_tasks);
  d_w21_ = upload(t.w21_tasks);
  d_pack_ = upload(t.pack_tasks);
  d_fwd_ = upload(t.fwd_tasks);
  d_bwd_ = upload(t.bwd_tasks);
  d_fwl_ = upload(t.fwl_tasks);
  d_bwl_ = upload(t.bwl_tasks);
  d_fwr_ = upload(t.fwr_tasks);
  d_bwr_ = upload(t.bwr_tasks);
  d_fws_ = upload(t.fws_tasks);
  d_bws_ = upload(t.bws_tasks);
  d_gat_all_ = upload(t.gat_all);
  d_gat_big_ = upload(t.gat_big);
  d_xg_all_ = upload(t.xg_all);
  d_xg_big_ = upload(t.xg_big);
  plan_ = t.plan;
  ++graph_epoch_;
  d_gr_ = dev<GT>(t.g_pool);
  d_gc_ = dev<GT>(t.g_pool);
  // the tiles' padding is read by the solves and never written: clear once
  HIPCHECK(hipMemset(d_gr_, 0, std::max<uint64_t>(t.g_pool, 1) * sizeof(GT)));
  HIPCHECK(hipMemset(d_gc_, 0, std::max<uint64_t>(t.g_pool, 1) * sizeof(GT)));
  d_body_reach_ = dev<double>(d_.B);
  d_body_inertia_ = dev<double>(d_.B);
  d_body_damp_ = dev<double>(d_.B);
  sym_ready_ = true;
}

// Numeric factorization (refreshFactor, factor.cc:567-613) of
// inertia M + reach sum R K R^T with the committed rotations.
// The tasks of launch range [s0, s1) for the bodies flagged in
// compact_mask_, appended to the pinned staging (their device copy lands at
// the same offset once staged); every task when no mask is set.
template <class Real>
template <class T>
typename Engine<Real>::TaskSpan Engine<Real>::compact(const std::vector<T> &tasks, const T *dev, uint32_t s0,
                                                      uint32_t s1) {
  if (!compact_mask_) return {dev + s0, s1 - s0};
  T *host;
  T *ddev;
  size_t *n;
  if constexpr (sizeof(T) == sizeof(uint64_t)) {
    host = hcompact64_;
    ddev = dcompact64_;
    n = &n64_;
  } else {
    host = hcompact32_;
    ddev = dcompact32_;
    n = &n32_;
  }
  const size_t begin = *n;
  const std::vector<int> &mask = *compact_mask_;
  const FactorTables &t = tab_;
  auto body_of = [&](uint32_t q) { return t.body[uint32_t(tasks[q])]; };
  uint32_t q = s0;
  while (q < s1) {
    const uint32_t b = body_of(q);
    // the end of body b's run
    uint32_t lo = q, hi = s1;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (body_of(mid) <= b) lo = mid + 1;
      else hi = mid;
    }
    if (mask[b]) {
      std::copy(tasks.begin() + q, tasks.begin() + lo, host + *n);
      *n += lo - q;
    }
    q = lo;
  }
  return {ddev + begin, uint32_t(*n - begin)};
}

template <class Real>
void Engine<Real>::factorize(const std::vector<double> &reach, const std::vector<double> &inertia,
                             const std::vector<double> *damp,
                             const std::vector<int> *mask, FactorProfile *prof) {
  if (!sym_ready_) analyze_factor();
  const std::vector<int> every(d_.B, 1);
  const std::vector<int> &fresh = mask ? *mask : every;
  hipStream_t s = stream_;
  hipEvent_t a, b;
  HIPCHECK(hipEventCreate(&a));
  HIPCHECK(hipEventCreate(&b));
  HIPCHECK(hipEventRecord(a, s));
  const FactorTables &t = tab_;
  HIPCHECK(hipMemcpyAsync(d_body_reach_, reach.data(), d_.B * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(d_body_inertia_, inertia.data(), d_.B * sizeof(double), hipMemcpyHostToDevice, s));
  // the damping coefficient of each body, 0 for the factor of the plain
  // operator (refreshFactor's damping argument, factor.cc:567-598)
  fac_damp_ = damp ? *damp : std::vector<double>(d_.B, 0.0);
  HIPCHECK(hipMemcpyAsync(d_body_damp_, fac_damp_.data(), d_.B * sizeof(double), hipMemcpyHostToDevice, s));
  FactorDamping<Real> damping;
  if (any_damped_) {
    damping.x = d_.x;
    damping.conn = d_.conn;
    damping.tau = d_.tau;
    damping.body_damp = d_body_damp_;
  }
  HIPCHECK(hipMemcpyAsync(d_refresh_, fresh.data(), d_.B * sizeof(int), hipMemcpyHostToDevice, s));
  // the launches' task ranges: a partial refresh stages compact lists of the
  // refreshed bodies' tasks and uploads them before the first launch
  compact_mask_ = mask && compactable_ && !prof ? mask : nullptr;
  n64_ = n32_ = 0;
  const TaskSpan zero = compact(t.zero_tasks, d_zero_, 0u, uint32_t(t.zero_tasks.size()));
  std::vector<TaskSpan> ea_span, potrf_span, trsm_span, syrk_span, trtri_span;
  for (size_t r = 0; r + 1 < t.ea_start.size(); ++r) ea_span.push_back(compact(t.ea_tasks, d_ea_, t.ea_start[r], t.ea_start[r + 1]));
  for (size_t r = 0; r + 1 < t.potrf_start.size(); ++r)
    potrf_span.push_back(compact(t.potrf_tasks, d_potrf_, t.potrf_start[r], t.potrf_start[r + 1]));
  for (size_t r = 0; r + 1 < t.trsm_start.size(); ++r)
    trsm_span.push_back(compact(t.trsm_tasks, d_trsm_, t.trsm_start[r], t.trsm_start[r + 1]));
  for (size_t r = 0; r + 1 < t.syrk_start.size(); ++r)
    syrk_span.push_back(compact(t.syrk_tasks, d_syrk_, t.syrk_start[r], t.syrk_start[r + 1]));
  for (size_t r = 0; r + 1 < t.trtri_start.size(); ++r)
    trtri_span.push_back(compact(t.trtri_tasks, d_trtri_, t.trtri_start[r], t.trtri_start[r + 1]));
  const TaskSpan w21 = compact(t.w21_tasks, d_w21_, 0u, uint32_t(t.w21_tasks.size()));
  const TaskSpan pack = compact(t.pack_tasks, d_pack_, 0u, uint32_t(t.pack_tasks.size()));
  if (compact_mask_) {
    if (n64_) HIPCHECK(hipMemcpyAsync(dcompact64_, hcompact64_, n64_ * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    if (n32_) HIPCHECK(hipMemcpyAsync(dcompact32_, hcompact32_, n32_ * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  }
  compact_mask_ = nullptr;
  auto u64 = [](const TaskSpan &x) { return static_cast<const uint64_t *>(x.dev); };
  if (mask) {
    // partial refresh: only the flagged bodies' fronts are rebuilt
    factor_zero(fd_, u64(zero), zero.count, s);
  } else {
    HIPCHECK(hipMemsetAsync(fd_.pool, 0, std::max<uint64_t>(t.pool, 1) * sizeof(double), s));
    HIPCHECK(hipMemsetAsync(fd_.minv, 0, std::max<uint64_t>(t.minv_pool, 1) * sizeof(double), s));
  }
  HIPCHECK(hipMemsetAsync(fd_.failed, 0, d_.B * sizeof(int), s));
  // every launch, with an event pair and its flops when profiled
  constexpr double TILE3 = 2.0 * FNB * FNB * FNB;  // one 64x64x64 MFMA tile
  auto dim = [](uint32_t total, uint32_t at) { return double(std::min<uint32_t>(FNB, total - at)); };
  auto timed = [&](uint32_t kind, uint64_t workgroups, auto &&flops, auto &&launch) {
    if (!prof || !workgroups) {
      launch();
      return;
    }
    typename FactorProfile::Launch L{kind, workgroups, 0.0, 0.0, nullptr, nullptr};
    flops(L.mfma, L.useful);
    HIPCHECK(hipEventCreate(&L.a));
    HIPCHECK(hipEventCreate(&L.b));
    HIPCHECK(hipEventRecord(L.a, s));
    launch();
    HIPCHECK(hipEventRecord(L.b, s));
    prof->launches.push_back(L);
  };
  auto none = [](double &, double &) {};
  timed(TALUS_MI_FK_ASSEMBLE, t.pair_target.size(), none, [&] {
    FactorLaunch<Real>::assemble(fd_, d_pair_target_, d_pair_ld_, d_pair_node_, d_pair_start_, d_contrib_,
                                 uint32_t(t.pair_target.size()), d_.dminv, d_.vol64, d_.lam64, d_.mu64, d_.R,
                                 d_.W, d_.plastic, d_.flags, d_.T, d_.mass, d_body_reach_, d_body_inertia_,
                                 d_pair_body_, damping, s);
  });
  size_t ea = 0, fac = 0;
  for (uint32_t l = 0; l < t.levels; ++l) {
    for (uint32_t r = 0; r < t.ea_rounds[l]; ++r, ++ea)
      timed(TALUS_MI_FK_EXTEND_ADD, t.ea_start[ea + 1] - t.ea_start[ea], none, [&] {
        factor_extend_add(fd_, d_parent_, u64(ea_span[ea]), ea_span[ea].count, s);
      });
    for (uint32_t p = 0; p < t.fac_panels[l]; ++p, ++fac) {
      const uint32_t P0 = p * FNB;
      timed(
          TALUS_MI_FK_POTRF, t.potrf_start[fac + 1] - t.potrf_start[fac],
          [&](double &, double &useful) {
            for (uint32_t q = t.potrf_start[fac]; q < t.potrf_start[fac + 1]; ++q) {
              const double nb = dim(t.k[t.potrf_tasks[q]], P0);
              useful += nb * nb * nb / 3.0;
            }
          },
          [&] {
            factor_potrf(fd_, static_cast<const uint32_t *>(potrf_span[fac].dev), potrf_span[fac].count, p, s);
          });
      timed(
          TALUS_MI_FK_TRSM, t.trsm_start[fac + 1] - t.trsm_start[fac],
          [&](double &mfma, double &useful) {
            for (uint32_t q = t.trsm_start[fac]; q < t.trsm_start[fac + 1]; ++q) {
              const uint64_t task = t.trsm_tasks[q];
              const uint32_t fr = uint32_t(task), rt = uint32_t((task >> 32) & 0xffffu);
              const double nb = dim(t.k[fr], P0);
              const double nr = dim(t.m[fr], P0 + uint32_t(nb) + rt * FNB);
              mfma += TILE3;
              useful += 2.0 * nr * nb * nb;
            }
          },
          [&] { factor_trsm(fd_, u64(trsm_span[fac]), trsm_span[fac].count, p, s); });
      timed(
          TALUS_MI_FK_SYRK, t.syrk_start[fac + 1] - t.syrk_start[fac],
          [&](double &mfma, double &useful) {
            for (uint32_t q = t.syrk_start[fac]; q < t.syrk_start[fac + 1]; ++q) {
              const uint64_t task = t.syrk_tasks[q];
              const uint32_t fr = uint32_t(task), I = uint32_t((task >> 32) & 0xffffu), J = uint32_t(task >> 48);
              const double nb = dim(t.k[fr], P0);
              const uint32_t R = P0 + uint32_t(nb);
              mfma += TILE3;
              useful += 2.0 * dim(t.m[fr], R + I * FNB) * dim(t.m[fr], R + J * FNB) * nb;
            }
          },
          [&] { factor_syrk(fd_, u64(syrk_span[fac]), syrk_span[fac].count, p, s); });
    }
  }
  for (size_t i = 0; i + 1 < t.trtri_start.size(); ++i)
    timed(
        TALUS_MI_FK_TRTRI, t.trtri_start[i + 1] - t.trtri_start[i],
        [&](double &mfma, double &useful) {
          for (uint32_t q = t.trtri_start[i]; q < t.trtri_start[i + 1]; ++q) {
            const uint64_t task = t.trtri_tasks[q];
            const uint32_t fr = uint32_t(task), bi = uint32_t((task >> 32) & 0xffffu), bj = uint32_t(task >> 48);
            const uint32_t k = t.k[fr];
            const double ni = dim(k, bi * FNB), nj = dim(k, bj * FNB);
            for (uint32_t c = bj; c < bi; ++c) useful += 2.0 * ni * dim(k, c * FNB) * nj;
            useful += ni * ni * nj;  // the scalar L_ii^-1 product
            mfma += double(bi - bj) * TILE3;
          }
        },
        [&] { factor_trtri(fd_, u64(trtri_span[i]), trtri_span[i].count, s); });
  timed(
      TALUS_MI_FK_PACK, t.w21_tasks.size() + t.pack_tasks.size(),
      [&](double &mfma, double &useful) {
        for (const uint64_t task : t.w21_tasks) {
          const uint32_t fr = uint32_t(task), I = uint32_t((task >> 32) & 0xffffu), J = uint32_t(task >> 48);
          const uint32_t k = t.k[fr], u = t.m[fr] - k, kt = (k + FNB - 1) / FNB;
          const double ni = dim(u, I * FNB), nj = dim(k, J * FNB);
          for (uint32_t c = J; c < kt; ++c) useful += 2.0 * ni * dim(k, c * FNB) * nj;
          mfma += double(kt - J) * TILE3;
        }
      },
      [&] {
        SolveLaunch<Real>::pack(fd_, u64(w21), w21.count, u64(pack), pack.count, d_gr_, d_gc_, s);
      });
  // the refreshed bodies' positions become their factor's reference frame
  Launch<Real>::frame_reference(d_, d_refresh_, s);
  HIPCHECK(hipMemcpyAsync(hfailed_, fd_.failed, d_.B * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipEventRecord(b, s));
  HIPCHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, a, b));
  HIPCHECK(hipEventDestroy(a));
  HIPCHECK(hipEventDestroy(b));
  factor_seconds_ = ms * 1e-3;
  fac_ok_ = true;
  for (uint32_t q = 0; q < d_.B; ++q) fac_ok_ = fac_ok_ && hfailed_[q] == 0;
  fac_ready_ = true;
  fac_reach_ = reach;
  fac_inertia_ = inertia;
  ++refactors_;
}

// One level's GEMVs with G in the plan's shape (0 rows per wavefront, 1 one
// row/column per lane, 2 64 rows/columns per workgroup split over its
// wavefronts; the fronts too large for 1 and 2 always use 0).
template <class Real>
void Engine<Real>::forward_level(uint32_t l, int mode, int parity, const Real *rin) {
  const FactorTables &t = tab_;
  if (!rin) rin = d_.r;
  if (mode == 1)
    SolveLaunch<Real>::forward_lanes(fd_, d_gc_, d_fwl_ + t.fwl_start[l], t.fwl_start[l + 1] - t.fwl_start[l], rin,
                                     d_.done, parity, d_.B, stream_);
  else if (mode == 2)
    SolveLaunch<Real>::forward_split(fd_, d_gc_, d_fws_ + t.fws_start[l], t.fws_start[l + 1] - t.fws_start[l], rin,
                                     d_.done, parity, d_.B, stream_);
  if (mode) {
    SolveLaunch<Real>::gather(fd_, d_gat_big_ + t.gat_big_start[l], t.gat_big_start[l + 1] - t.gat_big_start[l],
                              rin, d_.done, parity, d_.B, stream_);
    SolveLaunch<Real>::forward(fd_, d_gr_, d_fwr_ + t.fwr_start[l], t.fwr_start[l + 1] - t.fwr_start[l], rin,
                               d_.done, parity, d_.B, stream_);
  } else {
    SolveLaunch<Real>::gather(fd_, d_gat_all_ + t.gat_all_start[l], t.gat_all_start[l + 1] - t.gat_all_start[l],
                              rin, d_.done, parity, d_.B, stream_);
    SolveLaunch<Real>::forward(fd_, d_gr_, d_fwd_ + t.fwd_start[l], t.fwd_start[l + 1] - t.fwd_start[l], rin,
                               d_.done, parity, d_.B, stream_);
  }
}

template <class Real>
void Engine<Real>::backward_level(uint32_t l, int mode, int parity) {
  const FactorTables &t = tab_;
  if (mode == 1)
    SolveLaunch<Real>::backward_lanes(fd_, d_gr_, d_bwl_ + t.bwl_start[l], t.bwl_start[l + 1] - t.bwl_start[l],
                                      d_.done, parity, d_.B, stream_);
  else if (mode == 2)
    SolveLaunch<Real>::backward_split(fd_, d_gr_, d_bws_ + t.bws_start[l], t.bws_start[l + 1] - t.bws_start[l],
                                      d_.done, parity, d_.B, stream_);
  if (mode) {
    SolveLaunch<Real>::xgather(fd_, d_xg_big_ + t.xg_big_start[l], t.xg_big_start[l + 1] - t.xg_big_start[l],
                               d_.done, parity, d_.B, stream_);
    SolveLaunch<Real>::backward(fd_, d_gc_, d_bwr_ + t.bwr_start[l], t.bwr_start[l + 1] - t.bwr_start[l],
                                d_.done, parity, d_.B, stream_);
  } else {
    SolveLaunch<Real>::xgather(fd_, d_xg_all_ + t.xg_all_start[l], t.xg_all_start[l + 1] - t.xg_all_start[l],
                               d_.done, parity, d_.B, stream_);
    SolveLaunch<Real>::backward(fd_, d_gc_, d_bwd_ + t.bwd_start[l], t.bwd_start[l + 1] - t.bwd_start[l],
                                d_.done, parity, d_.B, stream_);
  }
}

// z = F (LL^T)^-1 F^T r for the bodies still iterating (applyFactor,
// factor.cc:648-704, F the body frame, or the identity): per level from the
// leaves, gather the right-hand side and the children's updates, then one
// GEMV with G; per level from the root, one transposed GEMV with G against
// [y; -x2], x2 read from the ancestors' solution at the front's update
// rows.  A fixed plan gives bit-identical results.
template <class Real>
void Engine<Real>::precondition(int parity) {
  hipStream_t s = stream_;
  const FactorTables &t = tab_;
  const uint32_t B = d_.B;
  // with bodyRotation the factor solves F^T r and the result turns by F
  if (frames_) Launch<Real>::frame_in(d_, parity, s);
  const Real *rin = frames_ ? d_.rf : d_.r;
  for (uint32_t l = 0; l < t.levels; ++l) forward_level(l, plan_[2 * l], parity, rin);
  for (uint32_t l = t.levels; l-- > 0;) backward_level(l, plan_[2 * l + 1], parity);
  SolveLaunch<Real>::finish(fd_, d_.z, d_.N, d_.anchored, d_node_body_, d_.done, parity, B,
                            frames_ ? d_.frame : nullptr, s);
}

// Autotuner: times both kernel shapes of every level's forward and backward
// pass on the device (HIP events, best of `repeats` after a warm-up) and
// keeps the faster; reports one preconditioner application before and after.
template <class Real>
void Engine<Real>::tune(uint32_t repeats, talus_mi_tune_report &rep) {
  std::memset(&rep, 0, sizeof rep);
  if (!built_ || d_.B == 0) return;
  if (!fac_ready_) {
    std::vector<double> reach(d_.B, 1.0), inert(d_.B, 1.0);
    factorize(reach, inert, nullptr);
    fac_reach_.clear();  // the next step refactors with its own keys
  }
  const uint32_t reps = std::max(repeats, 1u);
  hipStream_t s = stream_;
  hipEvent_t a, b;
  HIPCHECK(hipEventCreate(&a));
  HIPCHECK(hipEventCreate(&b));
  auto time = [&](auto &&launch) {
    launch();
    float best = 1e30f;
    for (uint32_t r = 0; r < reps; ++r) {
      HIPCHECK(hipEventRecord(a, s));
      launch();
      HIPCHECK(hipEventRecord(b, s));
      HIPCHECK(hipEventSynchronize(b));
      float ms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&ms, a, b));
      best = std::min(best, ms);
    }
    return double(best) * 1.0e-3;
  };
  const FactorTables &t = tab_;
  rep.levels = t.levels;
  frames_ = false;  // the solve kernels alone
  rep.before_seconds = time([&] { precondition(-1); });
  for (uint32_t l = 0; l < t.levels; ++l) {
    for (int dir = 0; dir < 2; ++dir) {
      const bool small = dir == 0 ? t.fwl_start[l + 1] > t.fwl_start[l] : t.bwl_start[l + 1] > t.bwl_start[l];
      int best = 0;
      if (small) {
        double fastest = 0.0;
        for (int mode = 0; mode < 3; ++mode) {
          const double sec = time([&] {
            if (dir == 0) forward_level(l, mode, -1);
            else backward_level(l, mode, -1);
          });
          if (mode == 0 || sec < fastest) {
            fastest = sec;
            best = mode;
          }
        }
      }
      plan_[2 * l + dir] = uint8_t(best);
    }
    rep.lanes_forward += plan_[2 * l] ? 1u : 0u;
    rep.lanes_backward += plan_[2 * l + 1] ? 1u : 0u;
  }
  ++graph_epoch_;
  rep.after_seconds = time([&] { precondition(-1); });
  HIPCHECK(hipEventDestroy(a));
  HIPCHECK(hipEventDestroy(b));
}

template <class Real>
void Engine<Real>::download(std::vector<Body> &bodies) {
  if (!built_) return;
  for (uint32_t b = 0; b < bodies.size() && b < n0_.size(); ++b) {
    Body &bd = bodies[b];
    if (bd.nodes != nn_[b] || bd.tets != tt_[b]) continue;
    bd.x.resize(3 * size_t(bd.nodes));
    bd.v.resize(3 * size_t(bd.nodes));
    bd.a.resize(3 * size_t(bd.nodes));
    state_get(b, bd.x.data(), bd.v.data(), bd.a.data());
    bd.R.resize(9 * size_t(bd.tets));
    rotations_get(b, bd.R.data());
    bd.W.resize(9 * size_t(bd.tets));
    frames_get(b, bd.W.data());
    bd.plastic.resize(10 * size_t(bd.tets));
    plastic_get(b, bd.plastic.data());
  }
}

template <class Real>
void Engine<Real>::upload_params(const std::vector<StepParams> &params, size_t slot) {
  StepParams *h = hparams_ + (slot % hparams_slots_) * d_.B;
  std::memcpy(h, params.data(), d_.B * sizeof(StepParams));
  HIPCHECK(hipMemcpyAsync(const_cast<StepParams *>(d_.params), h, d_.B * sizeof(StepParams),
                          hipMemcpyHostToDevice, stream_));
}

// parameters of one Newton operation from their own pinned ring; the ring
// restarts after every synchronization (nslot_ = 0), and waits for the
// stream when a round would reuse a slot still being copied
template <class Real>
void Engine<Real>::upload_newton(const std::vector<StepParams> &params) {
  if (nslot_ == NEWTON_SLOTS) {
    HIPCHECK(hipStreamSynchronize(stream_));
    nslot_ = 0;
  }
  StepParams *h = hnparams_ + nslot_ * d_.B;
  std::memcpy(h, params.data(), d_.B * sizeof(StepParams));
  HIPCHECK(hipMemcpyAsync(const_cast<StepParams *>(d_.params), h, d_.B * sizeof(StepParams),
                          hipMemcpyHostToDevice, stream_));
  ++nslot_;
}

// the per-solve result ring holds every solve of a step; grown (after a
// wait) when a Newton step needs more
template <class Real>
void Engine<Real>::ensure_solve_slot(uint32_t solve_index) {
  if (solve_index < hsolve_slots_) return;
  HIPCHECK(hipStreamSynchronize(stream_));
  nslot_ = 0;
  const size_t B = d_.B, grown = std::max<size_t>(2 * hsolve_slots_, solve_index + 1);
  uint32_t *it = pinned<uint32_t>(grown * B);
  double *rr = pinned<double>(grown * B), *tg = pinned<double>(grown * B);
  std::memcpy(it, hiters_, hsolve_slots_ * B * sizeof(uint32_t));
  std::memcpy(rr, hrr_, hsolve_slots_ * B * sizeof(double));
  std::memcpy(tg, htarget_, hsolve_slots_ * B * sizeof(double));
  hiters_ = it;
  hrr_ = rr;
  htarget_ = tg;
  hsolve_slots_ = grown;
}

// The plastic Newton of `bodies` for one substep pass, all bodies at once:
// each round collects every unfinished body's next operation, launches each
// kind of operation once for the bodies that asked for it (per-body
// parameters select them), waits, and hands every body its results.
template <class Real>
void Engine<Real>::newton_pass(const std::vector<uint32_t> &bodies, const std::vector<StepParams> &params,
                               uint32_t cap, uint32_t limit, bool diag, uint32_t pass,
                               std::vector<uint32_t> &solve_pass,
                               std::vector<std::vector<uint8_t>> &solve_active, uint32_t &solve_index,
                               talus_mi_report &rep) {
  hipStream_t s = stream_;
  const uint32_t B = d_.B;
  const size_t K = bodies.size();
  std::vector<NewtonBody> nb(K);
  std::vector<NewtonOp> op(K);
  for (size_t i = 0; i < K; ++i) {
    const StepParams &P = params[bodies[i]];
    nb[i].cap = cap;
    nb[i].travel = P.beta * P.inner * P.inner;
    op[i] = nb[i].next();
  }
  HIPCHECK(hipStreamSynchronize(s));
  nslot_ = 0;
  for (;;) {
    std::map<int, std::vector<size_t>> groups;
    for (size_t i = 0; i < K; ++i)
      if (op[i].kind != NewtonOp::DONE) groups[op[i].key()].push_back(i);
    if (groups.empty()) break;
    HIPCHECK(hipMemsetAsync(d_.loading, 0, B * sizeof(uint32_t), s));
    bool norms = false, loads = false;
    for (const auto &group : groups) {
      const NewtonOp o = op[group.second.front()];
      std::vector<StepParams> q = params;
      std::vector<uint8_t> mask(B, 0);
      for (StepParams &p : q) p.active = 0;
      for (size_t i : group.second) {
        StepParams &p = q[bodies[i]];
        p.active = 1;
        p.newton = 1;
        p.tangent = o.tangent;
        p.elastic_trial = o.elastic;
        p.margin = NewtonBody::MARGIN;
        mask[bodies[i]] = 1;
      }
      upload_newton(q);
      if (o.kind == NewtonOp::EVAL) {
        if (!o.first) Launch<Real>::evaluate(d_, s);
        Launch<Real>::assemble(d_, diag, true, s, o.first ? VISCOUS_PREDICTOR : VISCOUS_TAKEN);
        Launch<Real>::balance(d_, diag, !o.first, s);
        Launch<Real>::newton_norms(d_, s);
        norms = loads = true;
      } else if (o.kind == NewtonOp::SOLVE) {
        ensure_solve_slot(solve_index);
        solve_pass.push_back(pass);
        solve_active.push_back(mask);
        solve(o.first, limit, solve_index++);
        Launch<Real>::newton_norms(d_, s);
        norms = true;
      } else {
        Launch<Real>::newton_vec(d_, o.vec, s);
      }
    }
    if (norms)
      HIPCHECK(hipMemcpyAsync(hnnorm_, d_.nnorm, 3 * size_t(d_.n_ntiles) * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    if (loads) HIPCHECK(hipMemcpyAsync(hloading_, d_.loading, B * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    nslot_ = 0;
    for (size_t i = 0; i < K; ++i) {
      if (op[i].kind == NewtonOp::DONE) continue;
      const uint32_t b = bodies[i];
      if (op[i].kind == NewtonOp::EVAL) {
        double sq = 0.0;
        for (uint32_t t = nt0_[b]; t < nt1_[b]; ++t) sq += hnnorm_[3 * size_t(t)];
        nb[i].squared = sq;
        nb[i].loading = hloading_[b];
      } else if (op[i].kind == NewtonOp::SOLVE) {
        double moved = 0.0, extent = 0.0;
        for (uint32_t t = nt0_[b]; t < nt1_[b]; ++t) {
          moved = std::max(moved, hnnorm_[3 * size_t(t) + 1]);
          extent = std::max(extent, hnnorm_[3 * size_t(t) + 2]);
        }
        nb[i].moved = moved;
        nb[i].extent = extent;
      }
      op[i] = nb[i].next();
    }
  }
  for (size_t i = 0; i < nb.size(); ++i) {
    const NewtonBody &n = nb[i];
    talus_mi_report &own = body_rep_[bodies[i]];
    own.plastic_evaluations += n.evaluations;
    own.plastic_solves += n.solves_all;
    own.plastic_backtracks += n.backtracks;
    own.plastic_linear_reads += n.linear_reads;
    own.plastic_handoffs += n.handoffs;
    own.plastic_fallbacks += n.fallbacks;
    own.plastic_rejections += n.rejections;
    own.plastic_capped += n.capped;
    rep.plastic_evaluations += n.evaluations;
    rep.plastic_solves += n.solves_all;
    rep.plastic_backtracks += n.backtracks;
    rep.plastic_linear_reads += n.linear_reads;
    rep.plastic_handoffs += n.handoffs;
    rep.plastic_fallbacks += n.fallbacks;
    rep.plastic_rejections += n.rejections;
    rep.plastic_capped += n.capped;
  }
}

template <class Real>
void Engine<Real>::all_active() {
  std::vector<StepParams> p(d_.B);
  for (auto &q : p) {
    std::memset(&q, 0, sizeof q);
    q.active = 1;
  }
  upload_params(p, hparams_slots_ - 1);
}

// conjugateGradient for every active body at once.  The acceleration (first
// solve) or zero (corrections) is the starting iterate; the iterate change
// accumulates in dx and is added to a by apply_increment.
template <class Real>
void Engine<Real>::solve(bool first, uint32_t limit, uint32_t solve_index, bool tangent, double *rhs) {
  hipStream_t s = stream_;
  const bool fac = use_factor_;
  // tangent: the neo-Hookean Hessian operator of the adjoint solve, zero
  // initial guess, solution left in dx
  auto op = [&](const Real *dir, int parity) {
    if (tangent) Launch<Real>::hvp(d_, dir, parity, s);
    else Launch<Real>::op_element(d_, dir, parity, s);
  };
  if (first && !tangent) {
    op(d_.x0, -1);
    Launch<Real>::pcg_init(d_, true, fac, s);
  } else if (rhs) {
    Dev<Real> other = d_;
    other.right = rhs;
    Launch<Real>::pcg_init(other, false, fac, s);
  } else {
    Launch<Real>::pcg_init(d_, false, fac, s);
  }
  if (fac) {
    precondition(-1);
    ++applications_;
    Launch<Real>::pcg_rz(d_, true, -1, s);
  }
  Launch<Real>::pcg_begin(d_, s);
  // one PCG iteration: operator, gather, update, preconditioner, direction;
  // replayed from a captured HIP graph per (parity, preconditioner, operator)
  auto body = [&](int parity) {
    op(d_.p, parity);
    Launch<Real>::op_gather(d_, d_.p, d_.q, parity, s);
    Launch<Real>::pcg_update(d_, parity, fac, s);
    if (fac) {
      precondition(parity);
      Launch<Real>::pcg_rz(d_, false, parity, s);
    }
    Launch<Real>::pcg_direction(d_, parity, s);
  };
  auto iterate = [&](int parity) {
    if (!graphs_) {
      body(parity);
      return;
    }
    Graph &g = graph_[parity][fac ? 1 : 0][tangent ? 1 : 0];
    if (!g.exec || g.epoch != graph_epoch_) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
      hipGraph_t graph = nullptr;
      HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        body(parity);
      } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      HIPCHECK(hipStreamEndCapture(s, &graph));
      const hipError_t made = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHECK(made);
      g.epoch = graph_epoch_;
    }
    HIPCHECK(hipGraphLaunch(g.exec, s));
  };
  const uint32_t CHECK = fac ? 1u : 8u;
  for (uint32_t k = 0; k < limit; ++k) {
    const int parity = int(k & 1u);
    if (fac && k == 0) {
      // convergence of the starting residual is known only on the device
      HIPCHECK(hipMemcpyAsync(hflags_, d_.done, d_.B * sizeof(int), hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      bool all = true;
      for (uint32_t b = 0; b < d_.B && all; ++b) all = hflags_[b] != 0;
      if (all) break;
    }
    iterate(parity);
    if (fac) ++applications_;
    if ((k + 1) % CHECK == 0 && k + 1 < limit) {
      HIPCHECK(hipMemcpyAsync(hflags_, d_.done + size_t((k + 1) & 1u) * d_.B, d_.B * sizeof(int),
                              hipMemcpyDeviceToHost, s));
      HIPCHECK(hipStreamSynchronize(s));
      bool all = true;
      for (uint32_t b = 0; b < d_.B && all; ++b) all = hflags_[b] != 0;
      if (all) break;
    }
  }
  if (tangent) return;
  Launch<Real>::apply_increment(d_, first, s);
  const size_t slot = (solve_index % hsolve_slots_) * d_.B;
  HIPCHECK(hipMemcpyAsync(hiters_ + slot, d_.iters, d_.B * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(hrr_ + slot, d_.rr_last, d_.B * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(htarget_ + slot, d_.target2, d_.B * sizeof(double), hipMemcpyDeviceToHost, s));
}

// FP64 iterative refinement of the solve just made: the FP64 residual of
// every active body, then correction solves from zero on it (the PCG at
// the step's tolerance relative to the residual, so each round cuts the
// residual about by that factor) until every body meets refine_tolerance or
// refine_rounds corrections were made.
template <class Real>
void Engine<Real>::ensure_refine() {
  if (d_.slot64) return;
  const uint32_t B = d_.B;
  d_.slot64 = dev<double>(12 * size_t(d_.T));
  d_.rres = dev<double>(3 * size_t(d_.N));
  d_.rbase = dev<double>(3 * size_t(d_.N));
  d_.rpart = dev<double>(2 * size_t(d_.n_ntiles));
  d_.rrel = dev<double>(B);
  d_.rdone = dev<int>(B);
  hrdone_ = pinned<int>(B);
  hrrel_ = pinned<double>(B);
  ++graph_epoch_;  // captured iterations hold the old Dev
}

template <class Real>
void Engine<Real>::refine(const talus_mi_step_params &sp, uint32_t limit, uint32_t pass,
                          const std::vector<uint8_t> &mask, bool increment, std::vector<uint32_t> &solve_pass,
                          std::vector<std::vector<uint8_t>> &solve_active, uint32_t &solve_index,
                          RefineTally &tally) {
  hipStream_t s = stream_;
  const uint32_t B = d_.B;
  ensure_refine();
  if (tally.body_worst.size() != B) {
    tally.body_corrections.assign(B, 0);
    tally.body_unmet.assign(B, 0);
    tally.body_worst.assign(B, 0.0);
  }
  const uint32_t rounds = sp.refine_rounds;
  // a correction solve's unknown is its increment, kept as its own FP64 sum
  if (increment) Launch<Real>::track_increment(d_, false, s);
  for (uint32_t round = 0;; ++round) {
    Launch<Real>::refine_residual(d_, increment ? d_.rbase : nullptr, double(sp.refine_tolerance), s);
    HIPCHECK(hipMemcpyAsync(hrdone_, d_.rdone, B * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(hrrel_, d_.rrel, B * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    bool all = true;
    for (uint32_t b = 0; b < B && all; ++b) all = hrdone_[b] != 0;
    if (all || round == rounds) {
      for (uint32_t b = 0; b < B; ++b) {
        if (!mask.empty() && !mask[b]) continue;
        tally.worst = std::max(tally.worst, hrrel_[b]);
        tally.body_worst[b] = std::max(tally.body_worst[b], hrrel_[b]);
        if (!hrdone_[b]) {
          ++tally.unmet;
          ++tally.body_unmet[b];
        }
      }
      return;
    }
    for (uint32_t b = 0; b < B; ++b)
      if (!hrdone_[b]) {
        ++tally.corrections;
        ++tally.body_corrections[b];
      }
    solve_pass.push_back(pass);
    solve_active.push_back(mask);
    ensure_solve_slot(solve_index);
    solve(false, limit, solve_index++, false, d_.rres);
    if (increment) Launch<Real>::track_increment(d_, true, s);
  }
}

template <class Real>
void Engine<Real>::step(const talus_mi_step_params &sp, talus_mi_report &rep,
                        std::vector<talus_mi_event> &events, uint32_t step_index,
                        std::vector<uint8_t> &torn) {
  const uint32_t B = d_.B;
  hipStream_t s = stream_;
  hipEvent_t began, ended;
  HIPCHECK(hipEventCreate(&began));
  HIPCHECK(hipEventCreate(&ended));
  HIPCHECK(hipEventRecord(began, s));
  // integratorRule (solver.cc:531-545), effectiveSystem (554-562)
  double spectral = sp.spectral_radius;
  spectral = spectral < 0.0 ? 0.0 : (spectral > 1.0 ? 1.0 : spectral);
  const double am = (2.0 * spectral - 1.0) / (spectral + 1.0);
  const double af = spectral / (spectral + 1.0);
  const double gamma = 0.5 - am + af;
  const double beta = 0.25 * (1.0 - am + af) * (1.0 - am + af);
  const uint32_t configured = std::max(sp.substeps, 1u);
  const bool factor = sp.preconditioner == TALUS_MI_FACTOR;
  const uint32_t limit =
      factor && sp.factor_budget ? std::min(sp.factor_budget, sp.cg_iterations) : sp.cg_iterations;
  auto params_for = [&](uint32_t substeps, bool active) {
    StepParams p;
    std::memset(&p, 0, sizeof p);
    const double inner = double(sp.dt) / substeps;
    p.inner = inner;
    p.carry = (0.5 - beta) * inner * inner;
    const double heff = std::sqrt((1.0 - af) * beta) * inner;
    p.reach = heff * heff;
    p.inertia = 1.0 - am;
    p.am = am;
    p.af = af;
    p.beta = beta;
    p.gamma = gamma;
    for (int k = 0; k < 3; ++k) p.g[k] = double(sp.gravity[k]);
    p.tolerance = double(sp.tolerance);
    p.active = active ? 1u : 0u;
    p.limit = limit;
    p.damp = (1.0 - af) * gamma * inner;
    p.lag = (1.0 - af) * inner;
    return p;
  };
  // predictorInverts at the configured substep
  std::vector<StepParams> params(B);
  for (uint32_t b = 0; b < B; ++b) params[b] = params_for(configured, true);
  upload_params(params, 0);
  HIPCHECK(hipMemsetAsync(d_.inverts, 0, B * sizeof(int), s));
  HIPCHECK(hipMemsetAsync(d_.yields, 0, B * sizeof(uint32_t), s));
  HIPCHECK(hipMemsetAsync(d_.yflag, 0, std::max<uint32_t>(d_.T, 1), s));
  Launch<Real>::predictor_inverts(d_, s);
  HIPCHECK(hipMemcpyAsync(hflags_, d_.inverts, B * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  std::vector<uint32_t> substeps(B);
  uint32_t passes = 0;
  for (uint32_t b = 0; b < B; ++b) {
    substeps[b] = hflags_[b] ? std::max(configured, 4u) : configured;
    passes = std::max(passes, substeps[b]);
  }
  if (passes + 1 > hparams_slots_) {
    // grow the pinned parameter ring (only for unusual substep counts)
    HIPCHECK(hipStreamSynchronize(s));
    hparams_slots_ = passes + 2;
    hparams_ = pinned<StepParams>(hparams_slots_ * B);
  }
  const bool converge = (sp.flags & TALUS_MI_CONVERGE) != 0;
  graphs_ = (sp.flags & TALUS_MI_DIRECT_LAUNCH) == 0;
  // FP64 iterative refinement of the elastic solves (not of converged
  // rollouts, whose corrections already run to a fixed point)
  const bool refining = sp.refine_tolerance > 0.0f && !converge;
  RefineTally tally;
  const uint32_t reads = converge ? 64u : 3u;  // solves per pass at most
  if (reads * passes > hsolve_slots_) {
    HIPCHECK(hipStreamSynchronize(s));
    hsolve_slots_ = reads * passes;
    hiters_ = pinned<uint32_t>(hsolve_slots_ * B);
    hrr_ = pinned<double>(hsolve_slots_ * B);
    htarget_ = pinned<double>(hsolve_slots_ * B);
  }
  if (converge && !hnorms_) hnorms_ = pinned<double>(2 * size_t(d_.n_ntiles));
  if (converge && !d_norms_) d_norms_ = dev<double>(2 * size_t(d_.n_ntiles));
  std::vector<uint32_t> solve_pass;
  std::vector<std::vector<uint8_t>> solve_active;  // bodies of each solve (empty: every body in the pass)
  uint32_t solve_index = 0;
  // plastic Newton (plasticPasses > 1; not for converged rollouts)
  const uint32_t cap = sp.plastic_passes;
  const bool newton = cap > 1 && !converge &&
                      std::any_of(newton_body_.begin(), newton_body_.end(), [](uint8_t f) { return f != 0; });
  talus_mi_report newton_rep;
  std::memset(&newton_rep, 0, sizeof newton_rep);
  body_rep_.assign(B, talus_mi_report{});
  const uint32_t refactors_before = refactors_;
  applications_ = 0;
  const size_t n3 = 3 * size_t(d_.N);
  for (uint32_t pass = 0; pass < passes; ++pass) {
    for (uint32_t b = 0; b < B; ++b) params[b] = params_for(substeps[b], pass < substeps[b]);
    if (factor) {
      // refresh when missing, when the step or inertia keys change, or after
      // a capped solve (solver.cc:3639-3655, 3839-3899)
      std::vector<double> reach(B), inertia(B), damp(B);
      for (uint32_t b = 0; b < B; ++b) {
        const StepParams q = params_for(substeps[b], true);
        reach[b] = q.reach;
        inertia[b] = q.inertia;
        damp[b] = q.damp;
      }
      if (!fac_ready_ || reach != fac_reach_ || inertia != fac_inertia_ || (any_damped_ && damp != fac_damp_)) {
        HIPCHECK(hipStreamSynchronize(s));
        factorize(reach, inertia, &damp);
        std::fill(want_refresh_.begin(), want_refresh_.end(), 0);
      } else if (pass == 0 && std::any_of(want_refresh_.begin(), want_refresh_.end(), [](int w) { return w != 0; })) {
        // bodies whose solves were slow or capped last step: rebuild their
        // fronts only
        HIPCHECK(hipStreamSynchronize(s));
        factorize(reach, inertia, &damp, &want_refresh_);
        for (uint32_t b = 0; b < B; ++b)
          if (want_refresh_[b]) next_refresh_[b] = steps_taken_ + REFRESH_COOLDOWN;
        std::fill(want_refresh_.begin(), want_refresh_.end(), 0);
      }
    }
    use_factor_ = factor && fac_ok_;
    const bool diag = !use_factor_;
    upload_params(params, 1 + pass);
    // bodyRotation at the substep's start positions (solver.cc:3811-3813)
    frames_ = use_factor_ && !(sp.flags & TALUS_MI_FIXED_FRAME);
    if (frames_) Launch<Real>::frames(d_, s);
    double *record = nullptr;
    if (recorder_) {
      // x_n, v_n, a_n now, a' after the pass (adjoint replay)
      HIPCHECK(hipMalloc(&record, 4 * n3 * sizeof(double)));
      recorder_->buffers.push_back(record);
      recorder_->params.push_back(params);
      HIPCHECK(hipMemcpyAsync(record, d_.x, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHECK(hipMemcpyAsync(record + n3, d_.v, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHECK(hipMemcpyAsync(record + 2 * n3, d_.a, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    Launch<Real>::predict(d_, s);
    // bodies with yielding corotated lanes take the plastic Newton
    std::vector<uint32_t> nbodies;
    uint32_t in_pass = 0;
    for (uint32_t b = 0; b < B; ++b) {
      if (pass >= substeps[b]) continue;
      ++in_pass;
      if (newton && newton_body_[b]) nbodies.push_back(b);
    }
    std::vector<uint8_t> reg_mask;
    if (!nbodies.empty()) {
      std::vector<StepParams> reg = params;
      reg_mask.assign(B, 0);
      for (uint32_t b = 0; b < B; ++b) reg_mask[b] = pass < substeps[b] ? 1 : 0;
      for (uint32_t b : nbodies) {
        reg[b].active = 0;
        reg_mask[b] = 0;
      }
      upload_newton(reg);
    }
    if (nbodies.size() < in_pass) {
      Launch<Real>::assemble(d_, diag, false, s, VISCOUS_PREDICTOR);
      Launch<Real>::balance(d_, diag, false, s);
      solve_pass.push_back(pass);
      solve_active.push_back(reg_mask);
      ensure_solve_slot(solve_index);
      solve(true, limit, solve_index++);
      if (refining) refine(sp, limit, pass, reg_mask, false, solve_pass, solve_active, solve_index, tally);
      for (uint32_t read = 1; read < reads; ++read) {
        Launch<Real>::evaluate(d_, s);
        Launch<Real>::assemble(d_, diag, converge || read == 2, s, VISCOUS_TAKEN);
        Launch<Real>::balance(d_, diag, true, s);
        solve_pass.push_back(pass);
        solve_active.push_back(reg_mask);
        ensure_solve_slot(solve_index);
        solve(false, limit, solve_index++);
        if (refining) refine(sp, limit, pass, reg_mask, true, solve_pass, solve_active, solve_index, tally);
        if (converge) {
          // corrections until max |increment| <= 1e-12 max |a| in every body
          Launch<Real>::increment_norms(d_, d_norms_, s);
          HIPCHECK(hipMemcpyAsync(hnorms_, d_norms_, 2 * size_t(d_.n_ntiles) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
          HIPCHECK(hipStreamSynchronize(s));
          bool all = true;
          for (uint32_t b = 0; b < B && all; ++b) {
            if (pass >= substeps[b]) continue;
            double dx = 0.0, a = 0.0;
            for (uint32_t t = nt0_[b]; t < nt1_[b]; ++t) {
              dx = std::max(dx, hnorms_[2 * t]);
              a = std::max(a, hnorms_[2 * t + 1]);
            }
            all = dx <= 1.0e-12 * a;
          }
          if (all) break;
        }
      }
    }
    if (!nbodies.empty()) {
      newton_pass(nbodies, params, cap, limit, diag, pass, solve_pass, solve_active, solve_index, newton_rep);
      upload_newton(params);
    }
    Launch<Real>::finish(d_, s);
    if (any_plastic_) Launch<Real>::commit(d_, s);
    if (record)
      HIPCHECK(hipMemcpyAsync(record + 3 * n3, d_.a, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  // report reductions over every body
  for (uint32_t b = 0; b < B; ++b) params[b] = params_for(substeps[b], true);
  upload_params(params, 1 + passes);
  Launch<Real>::kinetic(d_, s);
  const bool tears = (sp.flags & TALUS_MI_DETECT_TEARS) && any_tear_;
  if (tears) Launch<Real>::tear(d_, s);
  HIPCHECK(hipEventRecord(ended, s));
  HIPCHECK(hipMemcpyAsync(hpartial_, d_.partial, size_t(d_.n_ntiles) * NPART * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(hepartial_, d_.epartial, d_.n_ttiles * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(hyields_, d_.yields, B * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  const bool record = (sp.flags & TALUS_MI_RECORD_EVENTS) != 0;
  std::vector<uint8_t> yflag, tflag;
  if (record && any_plastic_) {
    yflag.resize(d_.T);
    HIPCHECK(hipMemcpyAsync(yflag.data(), d_.yflag, d_.T, hipMemcpyDeviceToHost, s));
  }
  if (tears) {
    tflag.resize(d_.T);
    HIPCHECK(hipMemcpyAsync(tflag.data(), d_.tflag, d_.T, hipMemcpyDeviceToHost, s));
  }
  HIPCHECK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, began, ended));
  HIPCHECK(hipEventDestroy(began));
  HIPCHECK(hipEventDestroy(ended));
  // report (solver.h StepReport semantics, summed over bodies)
  std::memset(&rep, 0, sizeof rep);
  rep.bodies = B;
  rep.nodes = d_.N;
  rep.elements = d_.T;
  rep.substeps = passes;
  rep.refactors = refactors_ - refactors_before;
  rep.preconditioner_applications = applications_;
  rep.device_seconds = ms * 1.0e-3;
  std::vector<uint32_t> worst(B, 0);
  std::vector<int> capped_body(B, 0);
  for (uint32_t si = 0; si < solve_index; ++si) {
    const uint32_t pass = solve_pass[si];
    const size_t slot = (si % hsolve_slots_) * B;
    const std::vector<uint8_t> &mask = solve_active[si];
    for (uint32_t b = 0; b < B; ++b) {
      if (pass >= substeps[b]) continue;
      if (!mask.empty() && !mask[b]) continue;
      talus_mi_report &own = body_rep_[b];
      rep.cg_iterations += hiters_[slot + b];
      own.cg_iterations += hiters_[slot + b];
      worst[b] = std::max(worst[b], hiters_[slot + b]);
      const bool capped = hiters_[slot + b] >= limit && !(hrr_[slot + b] <= htarget_[slot + b]);
      capped_body[b] |= capped ? 1 : 0;
      rep.cg_capped += capped ? 1u : 0u;
      own.cg_capped += capped ? 1u : 0u;
      const double reference2 = htarget_[slot + b] / (sp.tolerance * double(sp.tolerance));
      const double rel = reference2 > 0.0 ? std::sqrt(hrr_[slot + b] / reference2) : 0.0;
      rep.cg_worst_relative = std::max(rep.cg_worst_relative, rel);
      own.cg_worst_relative = std::max(own.cg_worst_relative, rel);
      ++rep.solves;
      ++own.solves;
    }
  }
  for (uint32_t b = 0; b < B; ++b) {
    double kinetic = 0.0, bad = 0.0, elastic = 0.0;
    for (uint32_t t = nt0_[b]; t < nt1_[b]; ++t) {
      kinetic += hpartial_[t * NPART + 3];
      bad += hpartial_[t * NPART + 1];
    }
    for (uint32_t t = tt0_[b]; t < tt1_[b]; ++t) elastic += hepartial_[t];
    rep.kinetic_energy += kinetic;
    rep.elastic_energy += elastic;
    rep.non_finite |= bad > 0.0 ? 1u : 0u;
    rep.yield_events += hyields_[b];
    talus_mi_report &own = body_rep_[b];
    own.bodies = 1;
    own.nodes = nn_[b];
    own.elements = tt_[b];
    own.substeps = substeps[b];
    own.kinetic_energy = kinetic;
    own.elastic_energy = elastic;
    own.non_finite = bad > 0.0 ? 1u : 0u;
    own.yield_events = hyields_[b];
    own.plastic_growth = hyields_[b] ? 1u : 0u;
  }
  rep.plastic_growth = rep.yield_events ? 1u : 0u;
  rep.plastic_evaluations = newton_rep.plastic_evaluations;
  rep.plastic_solves = newton_rep.plastic_solves;
  rep.plastic_backtracks = newton_rep.plastic_backtracks;
  rep.plastic_linear_reads = newton_rep.plastic_linear_reads;
  rep.plastic_handoffs = newton_rep.plastic_handoffs;
  rep.plastic_fallbacks = newton_rep.plastic_fallbacks;
  rep.plastic_rejections = newton_rep.plastic_rejections;
  rep.plastic_capped = newton_rep.plastic_capped;
  rep.refinements = tally.corrections;
  rep.refine_unmet = tally.unmet;
  rep.refine_worst_relative = tally.worst;
  if (!tally.body_worst.empty())
    for (uint32_t b = 0; b < B; ++b) {
      body_rep_[b].refinements = tally.body_corrections[b];
      body_rep_[b].refine_unmet = tally.body_unmet[b];
      body_rep_[b].refine_worst_relative = tally.body_worst[b];
    }
  if (use_factor_) {
    for (uint32_t b = 0; b < B; ++b)
      if (capped_body[b] || (worst[b] >= REFRESH_ITERS && steps_taken_ >= next_refresh_[b]))
        want_refresh_[b] = 1;
  }
  ++steps_taken_;
  torn.assign(d_.T, 0);
  if (tears) torn = tflag;
  for (uint32_t b = 0; b < B; ++b) {
    for (uint32_t e = 0; e < tt_[b]; ++e) {
      const size_t g = t0_[b] + e;
      if (record && !yflag.empty() && yflag[g])
        events.push_back(talus_mi_event{TALUS_MI_EVENT_YIELD, b, e, step_index});
      if (tears && tflag[g]) {
        ++rep.tear_events;
        if (record) events.push_back(talus_mi_event{TALUS_MI_EVENT_TEAR, b, e, step_index});
      }
    }
  }
}

template <class Real>
void Engine<Real>::state_set(uint32_t b, const double *x, const double *v, const double *a) {
  const size_t off = 3 * size_t(n0_[b]), bytes = 3 * size_t(nn_[b]) * sizeof(double);
  if (x) HIPCHECK(hipMemcpy(d_.x + off, x, bytes, hipMemcpyHostToDevice));
  if (v) HIPCHECK(hipMemcpy(d_.v + off, v, bytes, hipMemcpyHostToDevice));
  if (a) HIPCHECK(hipMemcpy(d_.a + off, a, bytes, hipMemcpyHostToDevice));
}

template <class Real>
void Engine<Real>::state_get(uint32_t b, double *x, double *v, double *a) {
  HIPCHECK(hipStreamSynchronize(stream_));
  const size_t off = 3 * size_t(n0_[b]), bytes = 3 * size_t(nn_[b]) * sizeof(double);
  if (x) HIPCHECK(hipMemcpy(x, d_.x + off, bytes, hipMemcpyDeviceToHost));
  if (v) HIPCHECK(hipMemcpy(v, d_.v + off, bytes, hipMemcpyDeviceToHost));
  if (a) HIPCHECK(hipMemcpy(a, d_.a + off, bytes, hipMemcpyDeviceToHost));
}

template <class Real>
void Engine<Real>::planes_get(const Real *base, uint32_t b, double *out) {
  HIPCHECK(hipStreamSynchronize(stream_));
  const size_t T = d_.T, t0 = t0_[b], n = tt_[b];
  std::vector<Real> plane(n);
  for (int k = 0; k < 9; ++k) {
    HIPCHECK(hipMemcpy(plane.data(), base + k * T + t0, n * sizeof(Real), hipMemcpyDeviceToHost));
    for (size_t e = 0; e < n; ++e) out[9 * e + k] = double(plane[e]);
  }
}

template <class Real>
void Engine<Real>::planes_set(Real *base, uint32_t b, const double *in) {
  const size_t T = d_.T, t0 = t0_[b], n = tt_[b];
  std::vector<Real> plane(n);
  for (int k = 0; k < 9; ++k) {
    for (size_t e = 0; e < n; ++e) plane[e] = Real(in[9 * e + k]);
    HIPCHECK(hipMemcpy(base + k * T + t0, plane.data(), n * sizeof(Real), hipMemcpyHostToDevice));
  }
}

template <class Real>
void Engine<Real>::plastic_get(uint32_t b, double *p) {
  HIPCHECK(hipStreamSynchronize(stream_));
  const size_t T = d_.T, t0 = t0_[b], n = tt_[b];
  std::vector<double> plane(n);
  for (int k = 0; k < 10; ++k) {
    HIPCHECK(hipMemcpy(plane.data(), d_.plastic + k * T + t0, n * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t e = 0; e < n; ++e) p[10 * e + k] = plane[e];
  }
}

template <class Real>
void Engine<Real>::plastic_set(uint32_t b, const double *p) {
  const size_t T = d_.T, t0 = t0_[b], n = tt_[b];
  std::vector<double> plane(n);
  for (int k = 0; k < 10; ++k) {
    for (size_t e = 0; e < n; ++e) plane[e] = p[10 * e + k];
    HIPCHECK(hipMemcpy(d_.plastic + k * T + t0, plane.data(), n * sizeof(double), hipMemcpyHostToDevice));
  }
}

template <class Real>
void Engine<Real>::eval_forces(uint32_t b, const double *x, double *tf, double *te, double *nf,
                               double *R) {
  hipStream_t s = stream_;
  const size_t n0 = n0_[b], nn = nn_[b], t0 = t0_[b], nt = tt_[b];
  HIPCHECK(hipMemcpyAsync(scratch64_ + 3 * n0, x, 3 * nn * sizeof(double), hipMemcpyHostToDevice, s));
  all_active();
  Launch<Real>::displacement(d_, scratch64_, s);
  Launch<Real>::assemble(d_, false, true, s);
  Launch<Real>::gather_forces(d_, d_.q, s);
  // the body's slots are the CSR run of its nodes, in node order
  Real *h = hvec_;
  const size_t k0 = inc_start_[n0];
  HIPCHECK(hipMemcpyAsync(h, d_.slot_f + 3 * k0, 12 * nt * sizeof(Real), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(h + 12 * nt, d_.q + 3 * n0, 3 * nn * sizeof(Real), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  if (tf)
    for (size_t e = 0; e < nt; ++e)
      for (int c = 0; c < 4; ++c) {
        const size_t k = slot_pos_[4 * (t0 + e) + c] - k0;
        for (int r = 0; r < 3; ++r) tf[12 * e + 3 * c + r] = double(h[3 * k + r]);
      }
  if (nf)
    for (size_t i = 0; i < 3 * nn; ++i) nf[i] = double(h[12 * nt + i]);
  if (te) {
    std::vector<Real> en(nt);
    HIPCHECK(hipMemcpy(en.data(), d_.energy + t0, nt * sizeof(Real), hipMemcpyDeviceToHost));
    for (size_t e = 0; e < nt; ++e) te[e] = double(en[e]);
  }
  if (R) rotations_get(b, R);
}

// the body's operator with the same kernels as the allocation-free entry
// point, on the engine's FP64 staging and the body's own run of slots
template <class Real>
void Engine<Real>::apply_system(uint32_t b, double inertia, double step, const double *dv,
                                double *y) {
  hipStream_t s = stream_;
  const size_t n0 = n0_[b], nn = nn_[b];
  HIPCHECK(hipMemcpyAsync(scratch64_, dv, 3 * nn * sizeof(double), hipMemcpyHostToDevice, s));
  apply_system_device(b, inertia, step, scratch64_, scratch64_ + 3 * size_t(d_.N),
                      d_.slot_f + 3 * size_t(inc_start_[n0]), s);
  HIPCHECK(hipMemcpyAsync(y, scratch64_ + 3 * size_t(d_.N), 3 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
}

// the operator with the initial-stress term at positions x of body b: the
// element pass at x (rotations, corner forces), then the stressed operator
template <class Real>
void Engine<Real>::apply_system_stressed(uint32_t b, const double *x, double inertia, double step,
                                         const double *dv, double *y) {
  hipStream_t s = stream_;
  const size_t n0 = n0_[b], nn = nn_[b], N3 = 3 * size_t(d_.N);
  double *pos = scratch64_, *din = scratch64_ + N3, *yout = scratch64_ + 2 * N3;
  HIPCHECK(hipMemcpyAsync(pos, d_.x, N3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHECK(hipMemcpyAsync(pos + 3 * n0, x, 3 * nn * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(din, dv, 3 * nn * sizeof(double), hipMemcpyHostToDevice, s));
  all_active();
  Launch<Real>::displacement(d_, pos, s);
  Launch<Real>::assemble(d_, false, false, s);
  Launch<Real>::apply_body_stressed(d_, t0_[b], t0_[b] + tt_[b], uint32_t(n0), uint32_t(n0 + nn), inc_start_[n0],
                                    pos + 3 * n0, din, yout, inertia, step * step,
                                    d_.slot_f + 3 * size_t(inc_start_[n0]), s);
  HIPCHECK(hipMemcpyAsync(y, yout, 3 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
}

// the operator with the stiffness-damping term at positions x of body b:
// the element pass at x (rotations, rigid fits), then inertia M d +
// step^2 K d + damp sum_e tau_e C_e d
template <class Real>
void Engine<Real>::apply_system_damped(uint32_t b, const double *x, double inertia, double step, double damp,
                                       const double *dv, double *y) {
  hipStream_t s = stream_;
  const size_t n0 = n0_[b], nn = nn_[b], N3 = 3 * size_t(d_.N);
  double *pos = scratch64_, *din = scratch64_ + N3, *yout = scratch64_ + 2 * N3;
  HIPCHECK(hipMemcpyAsync(pos, d_.x, N3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHECK(hipMemcpyAsync(pos + 3 * n0, x, 3 * nn * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(din, dv, 3 * nn * sizeof(double), hipMemcpyHostToDevice, s));
  all_active();
  Launch<Real>::displacement(d_, pos, s);
  Launch<Real>::assemble(d_, false, false, s);
  Launch<Real>::apply_body_damped(d_, t0_[b], t0_[b] + tt_[b], uint32_t(n0), uint32_t(n0 + nn), inc_start_[n0],
                                  din, yout, inertia, step * step, damp,
                                  d_.slot_f + 3 * size_t(inc_start_[n0]), s);
  HIPCHECK(hipMemcpyAsync(y, yout, 3 * nn * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
}

template <class Real>
void Engine<Real>::time_kernel(uint32_t kind, uint32_t repeats, talus_mi_kernel_timing &out) {
  std::memset(&out, 0, sizeof out);
  hipStream_t s = stream_;
  const uint32_t reps = std::max(repeats, 1u);
  const double sz = double(sizeof(Real)), T = double(d_.T), N = double(d_.N);
  hipEvent_t a, b;
  HIPCHECK(hipEventCreate(&a));
  HIPCHECK(hipEventCreate(&b));
  if (kind <= 2) {
    all_active();
    // inputs at the current state
    Launch<Real>::displacement(d_, d_.x, s);
    HIPCHECK(hipMemsetAsync(d_.p, 0, 3 * size_t(d_.N) * sizeof(Real), s));
    auto launch = [&]() {
      if (kind == 0) Launch<Real>::assemble(d_, false, false, s);
      else if (kind == 1) Launch<Real>::op_element(d_, d_.p, -1, s);
      else Launch<Real>::op_gather(d_, d_.p, d_.q, -1, s);
    };
    launch();
    HIPCHECK(hipEventRecord(a, s));
    for (uint32_t i = 0; i < reps; ++i) launch();
    HIPCHECK(hipEventRecord(b, s));
    HIPCHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, a, b));
    out.avg_ms = double(ms) / reps;
    out.launches = 1;
    // element data once per tet (connectivity 16, slot positions 16, FP64
    // rest inverse 72, V/lambda/mu, flag byte), node vectors once per node
    const double tet = 16.0 + 16.0 + 72.0 + 3.0 * sz + 1.0;
    const double snh = double(n_snh_), plastic = double(n_plastic_);
    if (kind == 0)
      out.bytes = T * (tet + 9.0 * sz + 12.0 * sz) + snh * 9.0 * sz + plastic * (80.0 + 16.0) + N * 3.0 * sz;
    else if (kind == 1)  // the operator reads Dm^-1 in the element precision, the FP64
                         // plastic offsets Fp^-1 - I of plastic neo-Hookean tets, and
                         // its corners as 8-byte tile positions plus each tile's node
                         // list (4 bytes a node) instead of the 16-byte connectivity
      out.bytes = T * (tet - 72.0 - 16.0 + 8.0 + 9.0 * sz + 9.0 * sz + 12.0 * sz) + snh * 9.0 * sz +
                  double(n_flowing_) * 72.0 + 4.0 * double(tn_total_ + d_.n_ttiles + 1) + N * 3.0 * sz;
    else
      out.bytes = T * 12.0 * sz + (N + 1.0) * 4.0 + N * (6.0 * sz + 8.0 + 1.0);
  } else if (kind <= 8) {
    if (!fac_ready_) {
      std::vector<double> reach(d_.B, 1.0), inert(d_.B, 1.0);
      factorize(reach, inert, nullptr);
      fac_reach_.clear();  // the next step refactors with its own keys
    }
    const FactorTables &t = tab_;
    // kinds 3/4 lanes, 5/6 rows, 7/8 split; odd forward
    const bool forward = (kind % 2) == 1;
    const int shape = kind <= 4 ? 1 : (kind <= 6 ? 0 : 2);
    auto list = [&](uint32_t l) -> std::pair<const std::vector<uint64_t> *, const std::vector<uint32_t> *> {
      const int mode = plan_[2 * l + (forward ? 0 : 1)];
      if (shape == 0) {
        if (mode) return forward ? std::make_pair(&t.fwr_tasks, &t.fwr_start) : std::make_pair(&t.bwr_tasks, &t.bwr_start);
        return forward ? std::make_pair(&t.fwd_tasks, &t.fwd_start) : std::make_pair(&t.bwd_tasks, &t.bwd_start);
      }
      if (mode != shape) return {nullptr, nullptr};
      if (shape == 1) return forward ? std::make_pair(&t.fwl_tasks, &t.fwl_start) : std::make_pair(&t.bwl_tasks, &t.bwl_start);
      return forward ? std::make_pair(&t.fws_tasks, &t.fws_start) : std::make_pair(&t.bws_tasks, &t.bws_start);
    };
    // the levels this kind runs on under the current plan
    std::vector<uint32_t> levels;
    for (uint32_t l = 0; l < t.levels; ++l) {
      const auto [tasks, st] = list(l);
      if (tasks && (*st)[l + 1] > (*st)[l]) levels.push_back(l);
    }
    // algorithmic bytes: G entries on or below the diagonal once, the
    // front vector, the pivot outputs
    std::vector<uint8_t> in_kind(t.fronts, 0);
    for (uint32_t l : levels) {
      const auto [tasks, st] = list(l);
      for (uint32_t q = (*st)[l]; q < (*st)[l + 1]; ++q) in_kind[uint32_t((*tasks)[q] & 0xffffffffu)] = 1;
    }
    double bytes = 0.0;
    for (uint32_t f = 0; f < t.fronts; ++f) {
      if (!in_kind[f]) continue;
      const double k = t.k[f], m = t.m[f];
      const double g = m * k - 0.5 * k * (k - 1.0);
      // G in its storage type, the front vector, the column scales
      const double gsz = double(sizeof(GT));
      bytes += forward ? g * gsz + 16.0 * m + 8.0 * k : g * gsz + 8.0 * m + 20.0 * k;
    }
    all_active();
    auto run = [&](uint32_t l) {
      const auto [tasks, st] = list(l);
      const uint64_t *dev = nullptr;
      if (tasks == &t.fwl_tasks) dev = d_fwl_;
      else if (tasks == &t.bwl_tasks) dev = d_bwl_;
      else if (tasks == &t.fws_tasks) dev = d_fws_;
      else if (tasks == &t.bws_tasks) dev = d_bws_;
      else if (tasks == &t.fwr_tasks) dev = d_fwr_;
      else if (tasks == &t.bwr_tasks) dev = d_bwr_;
      else if (tasks == &t.fwd_tasks) dev = d_fwd_;
      else dev = d_bwd_;
      const uint32_t count = (*st)[l + 1] - (*st)[l];
      dev += (*st)[l];
      if (forward && shape == 1) SolveLaunch<Real>::forward_lanes(fd_, d_gc_, dev, count, d_.r, d_.done, -1, d_.B, s);
      else if (forward && shape == 2) SolveLaunch<Real>::forward_split(fd_, d_gc_, dev, count, d_.r, d_.done, -1, d_.B, s);
      else if (forward) SolveLaunch<Real>::forward(fd_, d_gr_, dev, count, d_.r, d_.done, -1, d_.B, s);
      else if (shape == 1) SolveLaunch<Real>::backward_lanes(fd_, d_gr_, dev, count, d_.done, -1, d_.B, s);
      else if (shape == 2) SolveLaunch<Real>::backward_split(fd_, d_gr_, dev, count, d_.done, -1, d_.B, s);
      else SolveLaunch<Real>::backward(fd_, d_gc_, dev, count, d_.done, -1, d_.B, s);
    };
    for (uint32_t l : levels) run(l);  // warm-up
    // one event pair per round of this kind's launches over its levels
    std::vector<hipEvent_t> ev(2 * reps);
    for (auto &x : ev) HIPCHECK(hipEventCreate(&x));
    for (uint32_t r = 0; r < reps; ++r) {
      HIPCHECK(hipEventRecord(ev[2 * r], s));
      for (uint32_t l : levels) run(l);
      HIPCHECK(hipEventRecord(ev[2 * r + 1], s));
    }
    HIPCHECK(hipStreamSynchronize(s));
    double total = 0.0;
    for (uint32_t r = 0; r < reps; ++r) {
      float ms = 0.0f;
      HIPCHECK(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]));
      total += ms;
    }
    for (auto &x : ev) HIPCHECK(hipEventDestroy(x));
    out.launches = uint32_t(levels.size());
    if (!levels.empty()) {
      out.avg_ms = total / double(reps * levels.size());
      out.bytes = bytes / double(levels.size());
    }
  }
  HIPCHECK(hipEventDestroy(a));
  HIPCHECK(hipEventDestroy(b));
}

template <class Real>
void Engine<Real>::factor_solve(uint32_t b, double inertia, double step, const double *r,
                                double *z) {
  hipStream_t s = stream_;
  std::vector<double> reach(d_.B, step * step), inert(d_.B, inertia);
  factorize(reach, inert, nullptr);
  frames_ = false;  // (LL^T)^-1 r of the factor just built
  const size_t n0 = n0_[b], nn = nn_[b];
  std::vector<Real> in(3 * size_t(d_.N), Real(0));
  for (size_t i = 0; i < 3 * nn; ++i) in[3 * n0 + i] = Real(r[i]);
  HIPCHECK(hipMemcpyAsync(d_.r, in.data(), in.size() * sizeof(Real), hipMemcpyHostToDevice, s));
  precondition(-1);
  std::vector<Real> out(3 * nn);
  HIPCHECK(hipMemcpyAsync(out.data(), d_.z + 3 * n0, 3 * nn * sizeof(Real), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  for (size_t i = 0; i < 3 * nn; ++i) z[i] = double(out[i]);
}

// Reverse pass of a converged rollout (implicit function theorem per
// substep; the derivation is in oracle/talus_oracle.c, adjoint_substep):
// per checkpoint segment from the last, replay the segment recording every
// substep, then walk the substeps backwards: x_w and the adjoint right-hand
// side, PCG on inertia M + h_eff^2 K(x_w) with the step's preconditioner,
// K lambda over every row, the adjoint state one substep back, and the
// parameter / rest-position gradients.
template <class Real>
void Engine<Real>::rollout_backward(const talus_mi_step_params &sp,
                                    const std::vector<std::vector<double>> &ckpt,
                                    const std::vector<double> &final_state, uint32_t steps,
                                    uint32_t every, uint32_t body, const double *x_bar,
                                    const double *v_bar, double tolerance, double *lambda_bar,
                                    double *mu_bar, double *rest_bar, double *x0_bar,
                                    double *v0_bar, double *a0_bar) {
  hipStream_t s = stream_;
  const size_t N = d_.N, T = d_.T, n3 = 3 * N;
  const size_t b0 = n0_[body], bn = nn_[body], e0 = t0_[body], et = tt_[body];
  std::vector<void *> mine;
  auto alloc = [&](size_t count) {
    double *p = nullptr;
    HIPCHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(double)));
    mine.push_back(p);
    HIPCHECK(hipMemsetAsync(p, 0, std::max<size_t>(count, 1) * sizeof(double), s));
    return p;
  };
  Recorder rec;
  try {
    AdjDev a{};
    a.xb = alloc(n3);
    a.vb = alloc(n3);
    a.ab = alloc(n3);
    a.kl = alloc(n3);
    a.w = alloc(N);
    a.lam_bar = alloc(T);
    a.mu_bar = alloc(T);
    a.xslot = alloc(12 * T);
    a.X_bar = alloc(n3);
    HIPCHECK(hipMemcpyAsync(a.xb + 3 * b0, x_bar, 3 * bn * sizeof(double), hipMemcpyHostToDevice, s));
    if (v_bar)
      HIPCHECK(hipMemcpyAsync(a.vb + 3 * b0, v_bar, 3 * bn * sizeof(double), hipMemcpyHostToDevice, s));
    talus_mi_step_params conv = sp;
    conv.flags = (sp.flags | TALUS_MI_CONVERGE) & ~(TALUS_MI_DETECT_TEARS | TALUS_MI_RECORD_EVENTS);
    const uint32_t limit = std::max(sp.cg_iterations, 2000u);
    std::vector<talus_mi_event> events;
    std::vector<uint8_t> torn;
    talus_mi_report rep;
    const uint32_t segments = (steps + every - 1) / every;
    for (uint32_t seg = segments; seg-- > 0;) {
      state_all_set(ckpt[seg]);
      rec = Recorder{};
      recorder_ = &rec;
      const uint32_t t1 = std::min(steps, (seg + 1) * every);
      for (uint32_t t = seg * every; t < t1; ++t) step(conv, rep, events, 0, torn);
      recorder_ = nullptr;
      for (size_t k = rec.buffers.size(); k-- > 0;) {
        std::vector<StepParams> params = rec.params[k];
        for (auto &p : params) {
          p.tolerance = tolerance;
          p.limit = limit;
        }
        HIPCHECK(hipStreamSynchronize(s));
        upload_params(params, 0);
        const double *r = rec.buffers[k];
        a.xn = r;
        a.vn = r + n3;
        a.an = r + 2 * n3;
        a.anext = r + 3 * n3;
        Launch<Real>::adj_prepare(d_, a, s);
        solve(false, limit, 0, true);
        Launch<Real>::hvp(d_, d_.dx, -1, s);
        Launch<Real>::adj_update(d_, a, s);
        Launch<Real>::adj_params(d_, a, s);
      }
      HIPCHECK(hipStreamSynchronize(s));
      for (double *p : rec.buffers) (void)hipFree(p);
      rec.buffers.clear();
    }
    HIPCHECK(hipStreamSynchronize(s));
    auto out = [&](double *dst, const double *src, size_t count) {
      if (dst) HIPCHECK(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyDeviceToHost));
    };
    out(lambda_bar, a.lam_bar + e0, et);
    out(mu_bar, a.mu_bar + e0, et);
    out(rest_bar, a.X_bar + 3 * b0, 3 * bn);
    out(x0_bar, a.xb + 3 * b0, 3 * bn);
    out(v0_bar, a.vb + 3 * b0, 3 * bn);
    out(a0_bar, a.ab + 3 * b0, 3 * bn);
    state_all_set(final_state);
  } catch (...) {
    recorder_ = nullptr;
    (void)hipStreamSynchronize(s);
    for (double *p : rec.buffers) (void)hipFree(p);
    for (void *p : mine) (void)hipFree(p);
    throw;
  }
  for (void *p : mine) (void)hipFree(p);
}

template <class Real>
void Engine<Real>::body_frame(uint32_t b, double *frame) {
  all_active();
  Launch<Real>::frames(d_, stream_);
  HIPCHECK(hipMemcpyAsync(frame, d_.frame + 9 * size_t(b), 9 * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIPCHECK(hipStreamSynchronize(stream_));
}

template <class Real>
void Engine<Real>::factor_profile(std::vector<talus_mi_factor_kernel> &out) {
  const std::vector<double> ones(d_.B, 1.0);
  const std::vector<double> reach = fac_reach_.size() == d_.B ? fac_reach_ : ones;
  const std::vector<double> inertia = fac_inertia_.size() == d_.B ? fac_inertia_ : ones;
  FactorProfile prof;
  const std::vector<double> damp = fac_damp_.size() == d_.B ? fac_damp_ : std::vector<double>(d_.B, 0.0);
  factorize(reach, inertia, &damp, nullptr, &prof);
  out.clear();
  for (const auto &L : prof.launches) {
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, L.a, L.b));
    (void)hipEventDestroy(L.a);
    (void)hipEventDestroy(L.b);
    auto it = std::find_if(out.begin(), out.end(), [&](const talus_mi_factor_kernel &k) { return k.kind == L.kind; });
    if (it == out.end()) {
      talus_mi_factor_kernel k;
      std::memset(&k, 0, sizeof k);
      k.kind = L.kind;
      out.push_back(k);
      it = out.end() - 1;
    }
    ++it->launches;
    it->workgroups += L.workgroups;
    it->mfma_flops += L.mfma;
    it->useful_flops += L.useful;
    it->ms += double(ms);
  }
}

template <class Real>
void Engine<Real>::factor_info(talus_mi_factor_info &info) {
  std::memset(&info, 0, sizeof info);
  if (!sym_ready_) return;
  info.fronts = tab_.fronts;
  info.levels = tab_.levels;
  info.ok = fac_ok_ ? 1u : 0u;
  info.entries = tab_.factor_entries;
  info.flops = tab_.flops;
  info.factor_seconds = factor_seconds_;
}


// Bodies split over several engines, each with its own HIP stream, stepped
// concurrently from one host thread per engine.  A batch's PCG runs in
// lockstep through the preconditioner's elimination levels, and the levels
// near the roots launch few workgroups; with two or three groups one group's
// small levels run beside another group's large ones.  Bodies go to the
// group with the fewest tets so far, largest body first (ties to the lower
// group), and keep their order inside a group.  Every per-body result is
// deterministic; results can differ from a one-group context in the last
// bits because each group's solve plan follows its own level structure.
class GroupedEngine final : public EngineBase {
 public:
  using Factory = std::function<std::unique_ptr<EngineBase>(hipStream_t)>;
  GroupedEngine(int device, uint32_t groups, hipStream_t first, const Factory &make) : device_(device) {
    streams_.push_back(first);
    for (uint32_t g = 1; g < groups; ++g) {
      hipStream_t s = nullptr;
      HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      streams_.push_back(s);
      owned_streams_.push_back(s);
    }
    for (uint32_t g = 0; g < groups; ++g) eng_.push_back(make(streams_[g]));
    sub_.resize(groups);
  }
  ~GroupedEngine() override {
    eng_.clear();
    for (hipStream_t s : owned_streams_) (void)hipStreamDestroy(s);
  }

  void build(const std::vector<Body> &bodies) override {
    const uint32_t G = uint32_t(eng_.size()), B = uint32_t(bodies.size());
    std::vector<uint32_t> order(B);
    for (uint32_t b = 0; b < B; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return bodies[a].tets > bodies[c].tets; });
    std::vector<uint64_t> load(G, 0);
    group_.assign(B, 0);
    for (uint32_t b : order) {
      const uint32_t g = uint32_t(std::min_element(load.begin(), load.end()) - load.begin());
      group_[b] = g;
      load[g] += bodies[b].tets;
    }
    members_.assign(G, {});
    local_.assign(B, 0);
    tet0_.assign(B, 0);
    for (uint32_t b = 0; b < B; ++b) {
      local_[b] = uint32_t(members_[group_[b]].size());
      members_[group_[b]].push_back(b);
      tet0_[b] = b ? tet0_[b - 1] + bodies[b - 1].tets : 0;
    }
    tets_ = B ? tet0_[B - 1] + bodies[B - 1].tets : 0;
    for (uint32_t g = 0; g < G; ++g) {
      sub_[g].clear();
      for (uint32_t b : members_[g]) sub_[g].push_back(bodies[b]);
      eng_[g]->build(sub_[g]);
    }
  }
  void download(std::vector<Body> &bodies) override {
    for (uint32_t g = 0; g < eng_.size(); ++g) {
      if (sub_[g].empty()) continue;
      eng_[g]->download(sub_[g]);
      for (uint32_t l = 0; l < members_[g].size(); ++l) {
        const uint32_t b = members_[g][l];
        if (b >= bodies.size()) continue;
        Body &dst = bodies[b];
        const Body &src = sub_[g][l];
        if (dst.nodes != src.nodes || dst.tets != src.tets) continue;
        dst.x = src.x;
        dst.v = src.v;
        dst.a = src.a;
        dst.R = src.R;
        dst.W = src.W;
        dst.plastic = src.plastic;
      }
    }
  }
  void step(const talus_mi_step_params &sp, talus_mi_report &rep, std::vector<talus_mi_event> &events,
            uint32_t step_index, std::vector<uint8_t> &torn) override {
    const uint32_t G = uint32_t(eng_.size());
    std::vector<talus_mi_report> reps(G);
    std::vector<std::vector<talus_mi_event>> evs(G);
    std::vector<std::vector<uint8_t>> torns(G);
    std::vector<std::exception_ptr> errs(G);
    // graph capture is stream-ordered state of the runtime; with several
    // threads capturing at once it failed, so the groups launch directly
    talus_mi_step_params direct = sp;
    direct.flags |= TALUS_MI_DIRECT_LAUNCH;
    auto work = [&](uint32_t g) {
      try {
        std::memset(&reps[g], 0, sizeof reps[g]);
        if (sub_[g].empty()) return;
        HIPCHECK(hipSetDevice(device_));
        eng_[g]->step(direct, reps[g], evs[g], step_index, torns[g]);
      } catch (...) {
        errs[g] = std::current_exception();
      }
    };
    std::vector<std::thread> threads;
    for (uint32_t g = 1; g < G; ++g) threads.emplace_back(work, g);
    work(0);
    for (std::thread &t : threads) t.join();
    for (const std::exception_ptr &e : errs)
      if (e) std::rethrow_exception(e);
    std::memset(&rep, 0, sizeof rep);
    for (const talus_mi_report &r : reps) {
      rep.bodies += r.bodies;
      rep.nodes += r.nodes;
      rep.elements += r.elements;
      rep.substeps = std::max(rep.substeps, r.substeps);
      rep.cg_iterations += r.cg_iterations;
      rep.cg_capped += r.cg_capped;
      rep.solves += r.solves;
      rep.refactors += r.refactors;
      rep.cg_worst_relative = std::max(rep.cg_worst_relative, r.cg_worst_relative);
      rep.elastic_energy += r.elastic_energy;
      rep.kinetic_energy += r.kinetic_energy;
      rep.plastic_growth |= r.plastic_growth;
      rep.yield_events += r.yield_events;
      rep.tear_events += r.tear_events;
      rep.non_finite |= r.non_finite;
      rep.preconditioner_applications += r.preconditioner_applications;
      rep.device_seconds = std::max(rep.device_seconds, r.device_seconds);
      rep.plastic_evaluations += r.plastic_evaluations;
      rep.plastic_solves += r.plastic_solves;
      rep.plastic_backtracks += r.plastic_backtracks;
      rep.plastic_linear_reads += r.plastic_linear_reads;
      rep.plastic_handoffs += r.plastic_handoffs;
      rep.plastic_fallbacks += r.plastic_fallbacks;
      rep.plastic_rejections += r.plastic_rejections;
      rep.plastic_capped += r.plastic_capped;
      rep.refinements += r.refinements;
      rep.refine_unmet += r.refine_unmet;
      rep.refine_worst_relative = std::max(rep.refine_worst_relative, r.refine_worst_relative);
    }
    // events in the one-engine order: by body, then tet, yield before tear
    const size_t first = events.size();
    for (uint32_t g = 0; g < G; ++g)
      for (talus_mi_event e : evs[g]) {
        e.body = members_[g][e.body];
        events.push_back(e);
      }
    std::stable_sort(events.begin() + first, events.end(),
                     [](const talus_mi_event &a, const talus_mi_event &c) { return a.body < c.body; });
    torn.assign(tets_, 0);
    for (uint32_t g = 0; g < G; ++g) {
      size_t off = 0;
      for (uint32_t b : members_[g]) {
        const uint32_t n = sub_[g][local_[b]].tets;
        if (torns[g].size() >= off + n)
          std::copy(torns[g].begin() + off, torns[g].begin() + off + n, torn.begin() + tet0_[b]);
        off += n;
      }
    }
  }
  void state_set(uint32_t b, const double *x, const double *v, const double *a) override {
    eng_[group_[b]]->state_set(local_[b], x, v, a);
  }
  void state_get(uint32_t b, double *x, double *v, double *a) override {
    eng_[group_[b]]->state_get(local_[b], x, v, a);
  }
  void rotations_get(uint32_t b, double *R) override { eng_[group_[b]]->rotations_get(local_[b], R); }
  void rotations_set(uint32_t b, const double *R) override { eng_[group_[b]]->rotations_set(local_[b], R); }
  void frames_get(uint32_t b, double *W) override { eng_[group_[b]]->frames_get(local_[b], W); }
  void frames_set(uint32_t b, const double *W) override { eng_[group_[b]]->frames_set(local_[b], W); }
  void plastic_get(uint32_t b, double *p) override { eng_[group_[b]]->plastic_get(local_[b], p); }
  void plastic_set(uint32_t b, const double *p) override { eng_[group_[b]]->plastic_set(local_[b], p); }
  void eval_forces(uint32_t b, const double *x, double *tf, double *te, double *nf, double *R) override {
    eng_[group_[b]]->eval_forces(local_[b], x, tf, te, nf, R);
  }
  void apply_system(uint32_t b, double inertia, double step, const double *dv, double *y) override {
    eng_[group_[b]]->apply_system(local_[b], inertia, step, dv, y);
  }
  void apply_system_device(uint32_t b, double inertia, double step, const double *dv, double *y, void *slots,
                           hipStream_t s) override {
    eng_[group_[b]]->apply_system_device(local_[b], inertia, step, dv, y, slots, s);
  }
  void apply_system_stressed(uint32_t b, const double *x, double inertia, double step, const double *dv,
                             double *y) override {
    eng_[group_[b]]->apply_system_stressed(local_[b], x, inertia, step, dv, y);
  }
  void apply_system_damped(uint32_t b, const double *x, double inertia, double step, double damp,
                           const double *dv, double *y) override {
    eng_[group_[b]]->apply_system_damped(local_[b], x, inertia, step, damp, dv, y);
  }
  void body_report(uint32_t b, talus_mi_report &out) const override {
    eng_[group_[b]]->body_report(local_[b], out);
  }
  void device_state(uint32_t b, const double **x, const double **v, hipStream_t *s) const override {
    eng_[group_[b]]->device_state(local_[b], x, v, s);
  }
  void factor_solve(uint32_t b, double inertia, double step, const double *r, double *z) override {
    eng_[group_[b]]->factor_solve(local_[b], inertia, step, r, z);
  }
  // per launch over every group's launches of the kind
  void time_kernel(uint32_t kind, uint32_t repeats, talus_mi_kernel_timing &out) override {
    std::memset(&out, 0, sizeof out);
    double time = 0.0, bytes = 0.0;
    for (uint32_t g = 0; g < eng_.size(); ++g) {
      if (sub_[g].empty()) continue;
      talus_mi_kernel_timing t;
      eng_[g]->time_kernel(kind, repeats, t);
      time += t.avg_ms * t.launches;
      bytes += t.bytes * t.launches;
      out.launches += t.launches;
    }
    if (out.launches) {
      out.avg_ms = time / out.launches;
      out.bytes = bytes / out.launches;
    }
  }
  void body_frame(uint32_t b, double *frame) override { eng_[group_[b]]->body_frame(local_[b], frame); }
  void factor_profile(std::vector<talus_mi_factor_kernel> &out) override {
    out.clear();
    for (auto &e : eng_) {
      std::vector<talus_mi_factor_kernel> one;
      e->factor_profile(one);
      for (const talus_mi_factor_kernel &k : one) {
        auto it = std::find_if(out.begin(), out.end(), [&](const talus_mi_factor_kernel &o) { return o.kind == k.kind; });
        if (it == out.end()) {
          out.push_back(k);
          continue;
        }
        it->launches += k.launches;
        it->workgroups += k.workgroups;
        it->mfma_flops += k.mfma_flops;
        it->useful_flops += k.useful_flops;
        it->ms += k.ms;
      }
    }
  }
  void factor_info(talus_mi_factor_info &info) override {
    std::memset(&info, 0, sizeof info);
    info.ok = 1;
    for (uint32_t g = 0; g < eng_.size(); ++g) {
      if (sub_[g].empty()) continue;
      talus_mi_factor_info i;
      eng_[g]->factor_info(i);
      info.fronts += i.fronts;
      info.levels = std::max(info.levels, i.levels);
      info.ok &= i.ok;
      info.entries += i.entries;
      info.flops += i.flops;
      info.factor_seconds += i.factor_seconds;
    }
  }
  void tune(uint32_t repeats, talus_mi_tune_report &report) override {
    std::memset(&report, 0, sizeof report);
    for (uint32_t g = 0; g < eng_.size(); ++g) {
      if (sub_[g].empty()) continue;
      talus_mi_tune_report r;
      eng_[g]->tune(repeats, r);
      report.levels += r.levels;
      report.lanes_forward += r.lanes_forward;
      report.lanes_backward += r.lanes_backward;
      report.before_seconds += r.before_seconds;
      report.after_seconds += r.after_seconds;
    }
  }
  void state_all_get(std::vector<double> &) override { unsupported(); }
  void state_all_set(const std::vector<double> &) override { unsupported(); }
  void rollout_backward(const talus_mi_step_params &, const std::vector<std::vector<double>> &,
                        const std::vector<double> &, uint32_t, uint32_t, uint32_t, const double *,
                        const double *, double, double *, double *, double *, double *, double *,
                        double *) override {
    unsupported();
  }
  // the groups' plans one after another
  void set_leaf(uint32_t leaf) override {
    for (auto &e : eng_) e->set_leaf(leaf);
  }
  std::vector<uint8_t> plan() override {
    std::vector<uint8_t> out;
    for (uint32_t g = 0; g < eng_.size(); ++g) {
      if (sub_[g].empty()) continue;
      const std::vector<uint8_t> p = eng_[g]->plan();
      out.insert(out.end(), p.begin(), p.end());
    }
    return out;
  }
  bool set_plan(const std::vector<uint8_t> &plan) override {
    std::vector<std::vector<uint8_t>> parts(eng_.size());
    size_t at = 0;
    for (uint32_t g = 0; g < eng_.size(); ++g) {
      if (sub_[g].empty()) continue;
      const size_t n = eng_[g]->plan().size();
      if (at + n > plan.size()) return false;
      parts[g].assign(plan.begin() + at, plan.begin() + at + n);
      at += n;
    }
    if (at != plan.size()) return false;
    for (size_t g = 0; g < eng_.size(); ++g)
      if (!sub_[g].empty() && !eng_[g]->set_plan(parts[g])) return false;
    return true;
  }

 private:
  [[noreturn]] static void unsupported() {
    throw std::runtime_error("differentiation needs a context with one stream (talus_mi_set_streams)");
  }
  int device_;
  std::vector<hipStream_t> streams_, owned_streams_;
  std::vector<std::unique_ptr<EngineBase>> eng_;
  std::vector<std::vector<Body>> sub_;
  std::vector<std::vector<uint32_t>> members_;
  std::vector<uint32_t> group_, local_;
  std::vector<size_t> tet0_;
  size_t tets_ = 0;
};

}  // namespace tmi

// ------------------------------------------------------------------ C ABI

struct talus_mi_ctx {
  int device = 0;
  uint32_t precision = TALUS_MI_FP64;
  hipStream_t stream = nullptr;
  std::vector<tmi::Body> bodies;
  std::unique_ptr<tmi::EngineBase> engine;
  bool dirty = true;
  uint64_t version = 0;  // counts edits that rebuild the engine
  // differentiable rollout: checkpoints (x, v, a of every body) and settings
  std::vector<std::vector<double>> ckpt;
  std::vector<double> ckpt_final;
  uint32_t ckpt_steps = 0, ckpt_every = 0;
  uint64_t ckpt_version = ~0ull;
  talus_mi_step_params ckpt_params{};
  uint32_t steps = 0;
  std::vector<talus_mi_event> events;
  uint32_t streams = 1;  // engines stepped concurrently (talus_mi_set_streams)
  uint32_t leaf = tmi::DEFAULT_LEAF;  // supernode leaf size (talus_mi_set_leaf)
  std::unique_ptr<tmi::LatticeSet> lattices;  // strainstep hex lattices
  // per lattice: the neo-Hookean hex terms and screen solve (built on use)
  std::vector<std::unique_ptr<tmi::HexScreen>> screens;
  // sampled fields resident for contact queries (talus_mi_field_add)
  std::vector<std::unique_ptr<tmi::DeviceField>> fields;
  // per body: its surface on the device for contact queries, and the
  // context version it was built at
  std::vector<std::unique_ptr<tmi::ContactFinder>> finders;
  std::vector<uint64_t> finder_version;
};

namespace {

using tmi::fail;

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[MAX_DEVICES] = {};

// resolves the array's address on `device`, which must be current; once per
// device (talus_mi_create)
void resolve_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("device index beyond the scratch table");
  if (g_scratch_ptr[device]) return;
  void *ptr = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (rc != hipSuccess || !ptr)
    throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
  g_scratch_ptr[device] = ptr;
}

void *scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return device >= 0 && device < MAX_DEVICES ? g_scratch_ptr[device] : nullptr;
}

// one engine on the context's stream, or bodies split over `streams` engines
std::unique_ptr<tmi::EngineBase> make_engine(talus_mi_ctx *ctx) {
  auto one = [precision = ctx->precision](hipStream_t s) -> std::unique_ptr<tmi::EngineBase> {
    if (precision == TALUS_MI_MIXED) return std::make_unique<tmi::Engine<float>>(s);
    return std::make_unique<tmi::Engine<double>>(s);
  };
  std::unique_ptr<tmi::EngineBase> e;
  if (ctx->streams <= 1) e = one(ctx->stream);
  else e = std::make_unique<tmi::GroupedEngine>(ctx->device, ctx->streams, ctx->stream, one);
  e->set_leaf(ctx->leaf);
  return e;
}

void ensure(talus_mi_ctx *ctx) {
  if (!ctx->dirty) return;
  HIPCHECK(hipSetDevice(ctx->device));
  ctx->engine->download(ctx->bodies);
  ctx->engine->build(ctx->bodies);
  ctx->dirty = false;
}

template <class F>
int guarded(talus_mi_ctx *ctx, F &&f) {
  if (!ctx) return fail(TALUS_MI_ERR_ARGUMENT, "null context");
  try {
    HIPCHECK(hipSetDevice(ctx->device));
    return f();
  } catch (const tmi::DeviceError &e) {
    return fail(TALUS_MI_ERR_DEVICE, e.what());
  } catch (const std::bad_alloc &) {
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory");
  } catch (const std::exception &e) {
    return fail(TALUS_MI_ERR_STATE, e.what());
  }
}

bool body_ok(talus_mi_ctx *ctx, uint32_t body) { return body < ctx->bodies.size(); }

}  // namespace

extern "C" {

const char *talus_mi_last_error(void) { return tmi::g_error.c_str(); }

void talus_mi_default_params(talus_mi_step_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->dt = 1.0f / 60.0f;
  p->gravity[1] = -9.81f;
  p->substeps = 1;
  p->cg_iterations = 64;
  p->factor_budget = 32;
  p->tolerance = 1.0e-3f;
  p->spectral_radius = 0.5f;
  p->preconditioner = TALUS_MI_BLOCK_JACOBI;
  p->flags = 0;
  p->plastic_passes = 40;
  p->refine_tolerance = 0.0f;
  p->refine_rounds = 8;
}

int talus_mi_create(talus_mi_ctx **out, int device, uint32_t precision) {
  if (!out) return fail(TALUS_MI_ERR_ARGUMENT, "null output pointer");
  *out = nullptr;
  if (precision != TALUS_MI_FP64 && precision != TALUS_MI_MIXED)
    return fail(TALUS_MI_ERR_ARGUMENT, "unknown precision");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return fail(TALUS_MI_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(TALUS_MI_ERR_ARGUMENT, "device index out of range");
  auto ctx = std::make_unique<talus_mi_ctx>();
  ctx->device = device;
  ctx->precision = precision;
  try {
    HIPCHECK(hipSetDevice(device));
    resolve_scratch(device);
    HIPCHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->engine = make_engine(ctx.get());
  } catch (const std::exception &e) {
    return fail(TALUS_MI_ERR_DEVICE, e.what());
  }
  *out = ctx.release();
  return TALUS_MI_OK;
}

int talus_mi_destroy(talus_mi_ctx *ctx) {
  if (!ctx) return TALUS_MI_OK;
  (void)hipSetDevice(ctx->device);
  ctx->engine.reset();
  ctx->lattices.reset();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return TALUS_MI_OK;
}

int talus_mi_build_lattice(const talus_mi_lattice *lattice, const talus_mi_material *materials,
                           uint32_t lanes, const talus_mi_support *supports, uint32_t nsupports,
                           talus_mi_mesh *out) {
  if (!lattice || !out) return fail(TALUS_MI_ERR_ARGUMENT, "null lattice or output");
  if (nsupports && !supports) return fail(TALUS_MI_ERR_ARGUMENT, "null supports");
  try {
    return tmi::build_lattice(lattice, materials, lanes, supports, nsupports, out);
  } catch (const std::bad_alloc &) {
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory");
  }
}

void talus_mi_mesh_free(talus_mi_mesh *mesh) { tmi::mesh_free(mesh); }

void talus_mi_free(void *p) { std::free(p); }

namespace {

int export_surface(const tmi::SurfaceMesh &sm, uint32_t tets, talus_mi_surface *out) {
  std::memset(out, 0, sizeof *out);
  out->triangles = uint32_t(sm.triangle_tet.size());
  out->edges = uint32_t(sm.edges.size() / 2);
  out->nodes = uint32_t(sm.nodes.size());
  out->tets = tets;
  auto copy = [](const auto &v) {
    using T = typename std::decay_t<decltype(v)>::value_type;
    T *p = (T *)std::malloc(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
  };
  out->triangle_nodes = copy(sm.triangles);
  out->triangle_tet = copy(sm.triangle_tet);
  out->triangle_lane = copy(sm.triangle_lane);
  out->edge_nodes = copy(sm.edges);
  out->surface_nodes = copy(sm.nodes);
  out->crack_faces = copy(sm.crack);
  out->clusters = sm.cluster_starts.empty() ? 0u : uint32_t(sm.cluster_starts.size() - 1);
  out->cluster_faces = copy(sm.cluster_faces);
  out->cluster_starts = copy(sm.cluster_starts);
  if (!out->triangle_nodes || !out->triangle_tet || !out->triangle_lane || !out->edge_nodes ||
      !out->surface_nodes || !out->crack_faces || !out->cluster_faces || !out->cluster_starts) {
    talus_mi_surface_free(out);
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory");
  }
  return TALUS_MI_OK;
}

}  // namespace

int talus_mi_mesh_surface(const talus_mi_mesh *mesh, talus_mi_surface *out) {
  if (!mesh || !out || !mesh->rest_positions || (mesh->tets && !mesh->tet_nodes))
    return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  try {
    for (size_t i = 0; i < 4 * size_t(mesh->tets); ++i)
      if (mesh->tet_nodes[i] >= mesh->nodes) return fail(TALUS_MI_ERR_ARGUMENT, "tet node out of range");
    std::vector<double> rest(mesh->rest_positions, mesh->rest_positions + 3 * size_t(mesh->nodes));
    tmi::SurfaceMesh sm;
    tmi::mesh_surface(mesh->nodes, mesh->tets, mesh->tet_nodes, mesh->tet_material, rest.data(), sm);
    return export_surface(sm, mesh->tets, out);
  } catch (const std::bad_alloc &) {
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory");
  }
}

int talus_mi_hex_mesh_surface(const talus_mi_hex_mesh *mesh, talus_mi_surface *out) {
  if (!mesh || !out || !mesh->rest_positions || (mesh->hexes && !mesh->hex_nodes))
    return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  try {
    for (size_t i = 0; i < 8 * size_t(mesh->hexes); ++i)
      if (mesh->hex_nodes[i] >= mesh->nodes) return fail(TALUS_MI_ERR_ARGUMENT, "hex node out of range");
    std::vector<double> rest(mesh->rest_positions, mesh->rest_positions + 3 * size_t(mesh->nodes));
    tmi::SurfaceMesh sm;
    tmi::hex_surface(mesh->nodes, mesh->hexes, mesh->hex_nodes, mesh->hex_material, rest.data(), sm);
    std::vector<uint8_t> on(mesh->nodes, 0);
    for (const uint32_t n : sm.triangles) on[n] = 1;
    for (uint32_t n = 0; n < mesh->nodes; ++n)
      if (on[n]) sm.nodes.push_back(n);
    sm.crack.assign(mesh->hexes, 0);
    return export_surface(sm, mesh->hexes, out);
  } catch (const std::bad_alloc &) {
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory");
  }
}

int talus_mi_lattice_surface_mesh(talus_mi_ctx *ctx, uint32_t lattice, talus_mi_surface *out) {
  if (!ctx || !out) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    if (!ctx->lattices || lattice >= ctx->lattices->count()) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    tmi::LatticeBody &b = ctx->lattices->body(lattice);
    b.ensure_surface();
    tmi::SurfaceMesh sm;
    tmi::hex_surface(b.nodes, b.hexes, b.hex.data(), b.lane.data(), b.rest.data(), sm);
    for (uint32_t n = 0; n < b.nodes; ++n)
      if (b.surface[n]) sm.nodes.push_back(n);
    sm.crack = b.crack.empty() ? std::vector<uint8_t>(b.hexes, 0) : b.crack;
    return export_surface(sm, b.hexes, out);
  });
}

int talus_mi_surface_get(talus_mi_ctx *ctx, uint32_t body, talus_mi_surface *out) {
  if (!ctx || !body_ok(ctx, body) || !out) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    const tmi::Body &b = ctx->bodies[body];
    tmi::SurfaceMesh sm;
    tmi::mesh_surface(b.nodes, b.tets, b.tet.data(), b.lane.data(), b.rest.data(), sm);
    return export_surface(sm, b.tets, out);
  });
}

int talus_mi_surface_thickness(const talus_mi_sdf *lattice, const double *rest, uint32_t triangles,
                               const uint32_t *triangle_nodes, float *thickness) {
  if (!lattice || !lattice->distances || (triangles && (!rest || !triangle_nodes || !thickness)))
    return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  for (int k = 0; k < 3; ++k)
    if (lattice->resolution[k] < 2u) return fail(TALUS_MI_ERR_ARGUMENT, "a sampled field needs two samples per axis");
  tmi::surface_thickness(*lattice, rest, triangles, triangle_nodes, thickness);
  return TALUS_MI_OK;
}

void talus_mi_surface_free(talus_mi_surface *s) {
  if (!s) return;
  std::free(s->triangle_nodes);
  std::free(s->triangle_tet);
  std::free(s->triangle_lane);
  std::free(s->edge_nodes);
  std::free(s->surface_nodes);
  std::free(s->crack_faces);
  std::free(s->cluster_faces);
  std::free(s->cluster_starts);
  std::memset(s, 0, sizeof *s);
}

int talus_mi_tear_mesh(const talus_mi_mesh *mesh, const talus_mi_material *materials,
                       uint32_t lanes, const double *x, const uint32_t *torn, uint32_t ntorn,
                       talus_mi_mesh *out, uint32_t **origin, uint32_t *new_nodes) {
  if (!mesh || !mesh->rest_positions || !mesh->tet_nodes || !mesh->anchored || !x || !out ||
      !origin || !new_nodes || (ntorn && !torn) || (lanes && !materials))
    return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  try {
    std::memset(out, 0, sizeof *out);
    *origin = nullptr;
    *new_nodes = 0;
    tmi::Body b;
    b.nodes = mesh->nodes;
    b.tets = mesh->tets;
    b.rest.resize(3 * size_t(b.nodes));
    for (size_t i = 0; i < b.rest.size(); ++i) b.rest[i] = double(mesh->rest_positions[i]);
    b.anchored.assign(mesh->anchored, mesh->anchored + b.nodes);
    b.tet.assign(mesh->tet_nodes, mesh->tet_nodes + 4 * size_t(b.tets));
    b.lane = mesh->tet_material ? std::vector<uint32_t>(mesh->tet_material, mesh->tet_material + b.tets)
                                : std::vector<uint32_t>(b.tets, 0u);
    b.materials.assign(materials, materials + lanes);
    int rc = tmi::prepare_body(b);
    if (rc != TALUS_MI_OK) return rc;
    b.x.assign(x, x + 3 * size_t(b.nodes));
    b.v.assign(3 * size_t(b.nodes), 0.0);
    b.a.assign(3 * size_t(b.nodes), 0.0);
    std::vector<uint32_t> list(torn, torn + ntorn), org;
    std::sort(list.begin(), list.end());
    list.erase(std::unique(list.begin(), list.end()), list.end());
    const uint32_t split = tmi::tear_topology(b, list, org);
    rc = tmi::prepare_body(b);
    if (rc != TALUS_MI_OK) return rc;
    out->nodes = b.nodes;
    out->tets = b.tets;
    out->rest_positions = (float *)std::malloc(3 * size_t(b.nodes) * sizeof(float));
    out->node_mass = (float *)std::malloc(size_t(b.nodes) * sizeof(float));
    out->anchored = (uint8_t *)std::malloc(b.nodes);
    out->tet_nodes = (uint32_t *)std::malloc(4 * size_t(b.tets) * sizeof(uint32_t));
    out->rest_inverse = (float *)std::malloc(9 * size_t(b.tets) * sizeof(float));
    out->volume = (float *)std::malloc(size_t(b.tets) * sizeof(float));
    out->tet_material = (uint32_t *)std::malloc(size_t(b.tets) * sizeof(uint32_t));
    *origin = (uint32_t *)std::malloc(std::max<size_t>(org.size(), 1) * sizeof(uint32_t));
    if (!out->rest_positions || !out->node_mass || !out->anchored || !out->tet_nodes ||
        !out->rest_inverse || !out->volume || !out->tet_material || !*origin) {
      tmi::mesh_free(out);
      std::free(*origin);
      *origin = nullptr;
      return fail(TALUS_MI_ERR_MEMORY, "out of host memory");
    }
    for (size_t i = 0; i < b.rest.size(); ++i) out->rest_positions[i] = float(b.rest[i]);
    for (uint32_t n = 0; n < b.nodes; ++n) out->node_mass[n] = float(b.mass[n]);
    std::memcpy(out->anchored, b.anchored.data(), b.nodes);
    std::memcpy(out->tet_nodes, b.tet.data(), b.tet.size() * sizeof(uint32_t));
    for (size_t i = 0; i < b.dminv.size(); ++i) out->rest_inverse[i] = float(b.dminv[i]);
    for (uint32_t e = 0; e < b.tets; ++e) out->volume[e] = float(b.volume[e]);
    std::memcpy(out->tet_material, b.lane.data(), b.lane.size() * sizeof(uint32_t));
    std::copy(org.begin(), org.end(), *origin);
    *new_nodes = split;
    return TALUS_MI_OK;
  } catch (const std::bad_alloc &) {
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory");
  }
}

int talus_mi_body_add(talus_mi_ctx *ctx, const talus_mi_mesh *mesh,
                      const talus_mi_material *materials, uint32_t lanes, uint32_t *body) {
  if (!mesh || !mesh->rest_positions || !mesh->tet_nodes || !mesh->anchored)
    return fail(TALUS_MI_ERR_ARGUMENT, "mesh is missing arrays");
  if (mesh->nodes == 0 || mesh->tets == 0) return fail(TALUS_MI_ERR_ARGUMENT, "empty mesh");
  if (lanes && !materials) return fail(TALUS_MI_ERR_ARGUMENT, "null materials");
  return guarded(ctx, [&]() -> int {
    tmi::Body b;
    b.nodes = mesh->nodes;
    b.tets = mesh->tets;
    b.rest.resize(3 * size_t(b.nodes));
    for (size_t i = 0; i < b.rest.size(); ++i) b.rest[i] = double(mesh->rest_positions[i]);
    b.anchored.assign(mesh->anchored, mesh->anchored + b.nodes);
    b.tet.assign(mesh->tet_nodes, mesh->tet_nodes + 4 * size_t(b.tets));
    if (mesh->tet_material)
      b.lane.assign(mesh->tet_material, mesh->tet_material + b.tets);
    else
      b.lane.assign(b.tets, 0u);
    b.materials.assign(materials, materials + lanes);
    b.density_mass = mesh->node_mass == nullptr;
    if (!b.density_mass) {
      b.mass.resize(b.nodes);
      for (uint32_t n = 0; n < b.nodes; ++n) b.mass[n] = double(mesh->node_mass[n]);
    }
    const int rc = tmi::prepare_body(b);
    if (rc != TALUS_MI_OK) return rc;
    b.x = b.rest;
    b.v.assign(3 * size_t(b.nodes), 0.0);
    b.a.assign(3 * size_t(b.nodes), 0.0);
    b.R.assign(9 * size_t(b.tets), 0.0);
    for (uint32_t e = 0; e < b.tets; ++e) b.R[9 * size_t(e)] = b.R[9 * size_t(e) + 4] = b.R[9 * size_t(e) + 8] = 1.0;
    b.W = b.R;
    b.plastic.assign(10 * size_t(b.tets), 0.0);
    ctx->engine->download(ctx->bodies);
    ctx->bodies.push_back(std::move(b));
    ctx->dirty = true;
    ++ctx->version;
    if (body) *body = uint32_t(ctx->bodies.size() - 1);
    return TALUS_MI_OK;
  });
}

int talus_mi_body_info(talus_mi_ctx *ctx, uint32_t body, uint32_t *nodes, uint32_t *tets) {
  if (!ctx || !body_ok(ctx, body)) return fail(TALUS_MI_ERR_ARGUMENT, "unknown body");
  if (nodes) *nodes = ctx->bodies[body].nodes;
  if (tets) *tets = ctx->bodies[body].tets;
  return TALUS_MI_OK;
}

int talus_mi_state_set(talus_mi_ctx *ctx, uint32_t body, const double *x, const double *v,
                       const double *a) {
  if (!ctx || !body_ok(ctx, body)) return fail(TALUS_MI_ERR_ARGUMENT, "unknown body");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->state_set(body, x, v, a);
    return TALUS_MI_OK;
  });
}

int talus_mi_state_get(talus_mi_ctx *ctx, uint32_t body, double *x, double *v, double *a) {
  if (!ctx || !body_ok(ctx, body)) return fail(TALUS_MI_ERR_ARGUMENT, "unknown body");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->state_get(body, x, v, a);
    return TALUS_MI_OK;
  });
}

int talus_mi_rotations_get(talus_mi_ctx *ctx, uint32_t body, double *R) {
  if (!ctx || !body_ok(ctx, body) || !R) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->rotations_get(body, R);
    return TALUS_MI_OK;
  });
}

int talus_mi_rotations_set(talus_mi_ctx *ctx, uint32_t body, const double *R) {
  if (!ctx || !body_ok(ctx, body) || !R) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->rotations_set(body, R);
    return TALUS_MI_OK;
  });
}

int talus_mi_frames_get(talus_mi_ctx *ctx, uint32_t body, double *W) {
  if (!ctx || !body_ok(ctx, body) || !W) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->frames_get(body, W);
    return TALUS_MI_OK;
  });
}

int talus_mi_frames_set(talus_mi_ctx *ctx, uint32_t body, const double *W) {
  if (!ctx || !body_ok(ctx, body) || !W) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->frames_set(body, W);
    return TALUS_MI_OK;
  });
}

int talus_mi_plastic_get(talus_mi_ctx *ctx, uint32_t body, double *p) {
  if (!ctx || !body_ok(ctx, body) || !p) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->plastic_get(body, p);
    return TALUS_MI_OK;
  });
}

int talus_mi_plastic_set(talus_mi_ctx *ctx, uint32_t body, const double *p) {
  if (!ctx || !body_ok(ctx, body) || !p) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->plastic_set(body, p);
    return TALUS_MI_OK;
  });
}

int talus_mi_topology_get(talus_mi_ctx *ctx, uint32_t body, uint32_t *tet_nodes, double *rest,
                          double *mass, uint8_t *anchored) {
  if (!ctx || !body_ok(ctx, body)) return fail(TALUS_MI_ERR_ARGUMENT, "unknown body");
  const tmi::Body &b = ctx->bodies[body];
  if (tet_nodes) std::memcpy(tet_nodes, b.tet.data(), b.tet.size() * sizeof(uint32_t));
  if (rest) std::memcpy(rest, b.rest.data(), b.rest.size() * sizeof(double));
  if (mass) std::memcpy(mass, b.mass.data(), b.mass.size() * sizeof(double));
  if (anchored) std::memcpy(anchored, b.anchored.data(), b.anchored.size());
  return TALUS_MI_OK;
}

int talus_mi_step(talus_mi_ctx *ctx, const talus_mi_step_params *params, talus_mi_report *report) {
  if (!params || !report) return fail(TALUS_MI_ERR_ARGUMENT, "null params or report");
  if (!(params->dt > 0.0f)) return fail(TALUS_MI_ERR_ARGUMENT, "dt must be positive");
  if (params->preconditioner > TALUS_MI_FACTOR) return fail(TALUS_MI_ERR_ARGUMENT, "unknown preconditioner");
  return guarded(ctx, [&]() -> int {
    std::memset(report, 0, sizeof *report);
    if (ctx->bodies.empty()) return TALUS_MI_OK;
    ensure(ctx);
    std::vector<uint8_t> torn;
    ctx->engine->step(*params, *report, ctx->events, ctx->steps, torn);
    ++ctx->steps;
    if ((params->flags & TALUS_MI_DETECT_TEARS) &&
        std::any_of(torn.begin(), torn.end(), [](uint8_t t) { return t != 0; })) {
      // host topology update after the step (overstretched + tearCells)
      ctx->engine->download(ctx->bodies);
      size_t t0 = 0;
      for (tmi::Body &b : ctx->bodies) {
        std::vector<uint32_t> list, origin;
        for (uint32_t e = 0; e < b.tets; ++e)
          if (torn[t0 + e]) list.push_back(e);
        t0 += b.tets;
        if (list.empty()) continue;
        const uint32_t split = tmi::tear_topology(b, list, origin);
        if (!split) continue;
        b.density_mass = true;
        const int rc = tmi::prepare_body(b);
        if (rc != TALUS_MI_OK) return rc;
        report->split_nodes += split;
        report->nodes += split;
        ctx->dirty = true;
        ++ctx->version;
      }
    }
    return TALUS_MI_OK;
  });
}

int talus_mi_body_report(talus_mi_ctx *ctx, uint32_t body, talus_mi_report *out) {
  if (!ctx || !body_ok(ctx, body) || !out) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    std::memset(out, 0, sizeof *out);
    if (ctx->engine && !ctx->dirty) ctx->engine->body_report(body, *out);
    return TALUS_MI_OK;
  });
}

int talus_mi_events(talus_mi_ctx *ctx, talus_mi_event *out, uint32_t capacity, uint32_t *count) {
  if (!ctx || !count) return fail(TALUS_MI_ERR_ARGUMENT, "null context or count");
  const uint32_t n = uint32_t(std::min<size_t>(capacity, ctx->events.size()));
  if (n && !out) return fail(TALUS_MI_ERR_ARGUMENT, "null output");
  std::copy(ctx->events.begin(), ctx->events.begin() + n, out);
  ctx->events.erase(ctx->events.begin(), ctx->events.begin() + n);
  *count = n;
  return TALUS_MI_OK;
}

int talus_mi_eval_forces(talus_mi_ctx *ctx, uint32_t body, const double *x, double *tet_forces,
                         double *tet_energy, double *node_forces, double *rotations) {
  if (!ctx || !body_ok(ctx, body) || !x) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->eval_forces(body, x, tet_forces, tet_energy, node_forces, rotations);
    return TALUS_MI_OK;
  });
}

int talus_mi_apply_system(talus_mi_ctx *ctx, uint32_t body, double inertia, double step,
                          const double *d, double *y) {
  if (!ctx || !body_ok(ctx, body) || !d || !y) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->apply_system(body, inertia, step, d, y);
    return TALUS_MI_OK;
  });
}

int talus_mi_apply_system_stressed(talus_mi_ctx *ctx, uint32_t body, const double *x, double inertia,
                                   double step, const double *d, double *y) {
  if (!ctx || !body_ok(ctx, body) || !x || !d || !y) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->apply_system_stressed(body, x, inertia, step, d, y);
    return TALUS_MI_OK;
  });
}

int talus_mi_apply_system_damped(talus_mi_ctx *ctx, uint32_t body, const double *x, double inertia,
                                 double step, double damp, const double *d, double *y) {
  if (!ctx || !body_ok(ctx, body) || !x || !d || !y) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->apply_system_damped(body, x, inertia, step, damp, d, y);
    return TALUS_MI_OK;
  });
}

int talus_mi_apply_system_noalloc(talus_mi_ctx *ctx, uint32_t body, double inertia, double step,
                                  const double *d, double *y) {
  if (!ctx || !body_ok(ctx, body) || !d || !y) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  const size_t real = ctx->precision == TALUS_MI_MIXED ? sizeof(float) : sizeof(double);
  const size_t need = 12 * size_t(ctx->bodies[body].tets) * real;
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(TALUS_MI_ERR_ARGUMENT, "body of " + std::to_string(ctx->bodies[body].tets) +
                                           " tets needs " + std::to_string(need) + " scratch bytes, above the " +
                                           std::to_string(NOALLOC_SCRATCH_BYTES) + " of the allocation-free path");
  if (ctx->dirty)
    return fail(TALUS_MI_ERR_STATE, "the context's device arrays are not built since the last body edit "
                                    "(call talus_mi_step or talus_mi_apply_system first)");
  return guarded(ctx, [&]() -> int {
    void *slots = scratch_of(ctx->device);
    if (!slots) return fail(TALUS_MI_ERR_STATE, "scratch address not resolved on this device");
    ctx->engine->apply_system_device(body, inertia, step, d, y, slots, ctx->stream);
    HIPCHECK(hipGetLastError());
    return TALUS_MI_OK;
  });
}

int talus_mi_time_kernel(talus_mi_ctx *ctx, uint32_t kind, uint32_t repeats,
                         talus_mi_kernel_timing *out) {
  if (!ctx || !out || kind > 8) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    if (ctx->bodies.empty()) return fail(TALUS_MI_ERR_STATE, "no bodies");
    ensure(ctx);
    ctx->engine->time_kernel(kind, repeats, *out);
    return TALUS_MI_OK;
  });
}

int talus_mi_factor_solve(talus_mi_ctx *ctx, uint32_t body, double inertia, double step,
                          const double *r, double *z) {
  if (!ctx || !body_ok(ctx, body) || !r || !z) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->factor_solve(body, inertia, step, r, z);
    return TALUS_MI_OK;
  });
}

int talus_mi_body_frame_get(talus_mi_ctx *ctx, uint32_t body, double *frame) {
  if (!ctx || !body_ok(ctx, body) || !frame) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->body_frame(body, frame);
    return TALUS_MI_OK;
  });
}

int talus_mi_factor_profile(talus_mi_ctx *ctx, talus_mi_factor_kernel *out, uint32_t capacity,
                            uint32_t *count) {
  if (!ctx || !count || (capacity && !out)) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    if (ctx->bodies.empty()) return fail(TALUS_MI_ERR_STATE, "no bodies");
    std::vector<talus_mi_factor_kernel> all;
    ctx->engine->factor_profile(all);
    *count = uint32_t(all.size());
    for (size_t i = 0; i < all.size() && i < capacity; ++i) out[i] = all[i];
    return TALUS_MI_OK;
  });
}

int talus_mi_factor_info_get(talus_mi_ctx *ctx, talus_mi_factor_info *info) {
  if (!info) return fail(TALUS_MI_ERR_ARGUMENT, "null info");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->factor_info(*info);
    return TALUS_MI_OK;
  });
}

int talus_mi_tune(talus_mi_ctx *ctx, uint32_t repeats, talus_mi_tune_report *report) {
  if (!ctx || !report) return fail(TALUS_MI_ERR_ARGUMENT, "null context or report");
  return guarded(ctx, [&]() -> int {
    std::memset(report, 0, sizeof *report);
    if (ctx->bodies.empty()) return TALUS_MI_OK;
    ensure(ctx);
    ctx->engine->tune(repeats, *report);
    report->leaf = ctx->leaf;
    return TALUS_MI_OK;
  });
}

int talus_mi_plan_get(talus_mi_ctx *ctx, uint8_t *plan, uint32_t capacity, uint32_t *count) {
  if (!ctx || !count) return fail(TALUS_MI_ERR_ARGUMENT, "null context or count");
  return guarded(ctx, [&]() -> int {
    *count = 0;
    if (ctx->bodies.empty()) return TALUS_MI_OK;
    ensure(ctx);
    const std::vector<uint8_t> p = ctx->engine->plan();
    *count = uint32_t(p.size());
    if (capacity < p.size()) return plan ? fail(TALUS_MI_ERR_ARGUMENT, "plan capacity too small") : TALUS_MI_OK;
    if (!plan) return fail(TALUS_MI_ERR_ARGUMENT, "null plan");
    std::copy(p.begin(), p.end(), plan);
    return TALUS_MI_OK;
  });
}

int talus_mi_plan_set(talus_mi_ctx *ctx, const uint8_t *plan, uint32_t count) {
  if (!ctx || (count && !plan)) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    if (ctx->bodies.empty()) return count ? fail(TALUS_MI_ERR_ARGUMENT, "no bodies") : TALUS_MI_OK;
    ensure(ctx);
    if (!ctx->engine->set_plan(std::vector<uint8_t>(plan, plan + count)))
      return fail(TALUS_MI_ERR_ARGUMENT, "plan does not match the factor's levels");
    return TALUS_MI_OK;
  });
}

void *talus_mi_stream(talus_mi_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int talus_mi_set_streams(talus_mi_ctx *ctx, uint32_t streams) {
  if (streams < 1 || streams > TALUS_MI_MAX_STREAMS)
    return fail(TALUS_MI_ERR_ARGUMENT, "streams must be 1.." + std::to_string(TALUS_MI_MAX_STREAMS));
  return guarded(ctx, [&]() -> int {
    if (streams == ctx->streams) return TALUS_MI_OK;
    if (!ctx->dirty) ctx->engine->download(ctx->bodies);  // keep the state of the current engine
    ctx->engine.reset();
    ctx->streams = streams;
    ctx->engine = make_engine(ctx);
    ctx->dirty = true;
    ++ctx->version;
    return TALUS_MI_OK;
  });
}

int talus_mi_set_leaf(talus_mi_ctx *ctx, uint32_t leaf) {
  if (leaf < TALUS_MI_LEAF_MIN || leaf > TALUS_MI_LEAF_MAX)
    return fail(TALUS_MI_ERR_ARGUMENT, "leaf must be " + std::to_string(TALUS_MI_LEAF_MIN) + ".." +
                                           std::to_string(TALUS_MI_LEAF_MAX));
  return guarded(ctx, [&]() -> int {
    if (leaf == ctx->leaf) return TALUS_MI_OK;
    if (!ctx->dirty) ctx->engine->download(ctx->bodies);  // keep the state of the current engine
    ctx->leaf = leaf;
    ctx->engine->set_leaf(leaf);
    ctx->dirty = true;
    ++ctx->version;
    return TALUS_MI_OK;
  });
}

int talus_mi_leaf_get(talus_mi_ctx *ctx, uint32_t *leaf) {
  if (!ctx || !leaf) return fail(TALUS_MI_ERR_ARGUMENT, "null argument");
  *leaf = ctx->leaf;
  return TALUS_MI_OK;
}

int talus_mi_tune_leaf(talus_mi_ctx *ctx, const uint32_t *candidates, uint32_t count, uint32_t repeats,
                       double *times, talus_mi_tune_report *report) {
  if (!ctx || !report || (count && !candidates)) return fail(TALUS_MI_ERR_ARGUMENT, "null argument");
  for (uint32_t i = 0; i < count; ++i)
    if (candidates[i] < TALUS_MI_LEAF_MIN || candidates[i] > TALUS_MI_LEAF_MAX)
      return fail(TALUS_MI_ERR_ARGUMENT, "leaf candidate out of range");
  std::memset(report, 0, sizeof *report);
  if (ctx->bodies.empty() || count == 0) return TALUS_MI_OK;
  double before = 0.0;
  int rc = talus_mi_tune(ctx, repeats, report);
  if (rc != TALUS_MI_OK) return rc;
  before = report->before_seconds;
  uint32_t best = ctx->leaf;
  double fastest = report->after_seconds;
  for (uint32_t i = 0; i < count; ++i) {
    if ((rc = talus_mi_set_leaf(ctx, candidates[i])) != TALUS_MI_OK) return rc;
    if ((rc = talus_mi_tune(ctx, repeats, report)) != TALUS_MI_OK) return rc;
    if (times) times[i] = report->after_seconds;
    if (report->after_seconds < fastest) {
      fastest = report->after_seconds;
      best = candidates[i];
    }
  }
  if ((rc = talus_mi_set_leaf(ctx, best)) != TALUS_MI_OK) return rc;
  if ((rc = talus_mi_tune(ctx, repeats, report)) != TALUS_MI_OK) return rc;
  report->before_seconds = before;
  return TALUS_MI_OK;
}

}  // extern "C"

// ------------------------------------------------------- differentiation

namespace {

int edit_body(talus_mi_ctx *ctx, uint32_t body, const std::function<int(tmi::Body &)> &edit) {
  if (!ctx || !body_ok(ctx, body)) return fail(TALUS_MI_ERR_ARGUMENT, "unknown body");
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    ctx->engine->download(ctx->bodies);
    tmi::Body &b = ctx->bodies[body];
    tmi::Body saved = b;
    int rc = edit(b);
    if (rc == TALUS_MI_OK) rc = tmi::prepare_body(b);
    if (rc != TALUS_MI_OK) {
      b = saved;
      return rc;
    }
    ctx->dirty = true;
    ++ctx->version;
    return TALUS_MI_OK;
  });
}

}  // namespace

int talus_mi_lame_set(talus_mi_ctx *ctx, uint32_t body, const double *lambda, const double *mu) {
  if (!lambda || !mu) return fail(TALUS_MI_ERR_ARGUMENT, "null parameters");
  return edit_body(ctx, body, [&](tmi::Body &b) -> int {
    for (uint32_t e = 0; e < b.tets; ++e)
      if (!(std::isfinite(lambda[e]) && std::isfinite(mu[e]) && mu[e] > 0.0 && lambda[e] >= 0.0))
        return fail(TALUS_MI_ERR_ARGUMENT, "Lame parameters must be finite, mu > 0, lambda >= 0");
    b.lambda.assign(lambda, lambda + b.tets);
    b.mu.assign(mu, mu + b.tets);
    b.lame_set = true;
    return TALUS_MI_OK;
  });
}

int talus_mi_lame_get(talus_mi_ctx *ctx, uint32_t body, double *lambda, double *mu) {
  if (!ctx || !body_ok(ctx, body) || !lambda || !mu) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  const tmi::Body &b = ctx->bodies[body];
  std::copy(b.lambda.begin(), b.lambda.end(), lambda);
  std::copy(b.mu.begin(), b.mu.end(), mu);
  return TALUS_MI_OK;
}

int talus_mi_rest_set(talus_mi_ctx *ctx, uint32_t body, const double *rest) {
  if (!rest) return fail(TALUS_MI_ERR_ARGUMENT, "null rest positions");
  return edit_body(ctx, body, [&](tmi::Body &b) -> int {
    b.rest.assign(rest, rest + 3 * size_t(b.nodes));
    return TALUS_MI_OK;
  });
}

int talus_mi_rollout(talus_mi_ctx *ctx, const talus_mi_step_params *params, uint32_t steps,
                     uint32_t every, talus_mi_report *report) {
  if (!params || !report || steps == 0) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  if (!(params->dt > 0.0f)) return fail(TALUS_MI_ERR_ARGUMENT, "dt must be positive");
  if (params->preconditioner > TALUS_MI_FACTOR) return fail(TALUS_MI_ERR_ARGUMENT, "unknown preconditioner");
  return guarded(ctx, [&]() -> int {
    std::memset(report, 0, sizeof *report);
    if (ctx->bodies.empty()) return fail(TALUS_MI_ERR_STATE, "no bodies");
    ensure(ctx);
    talus_mi_step_params p = *params;
    p.flags = (p.flags | TALUS_MI_CONVERGE) & ~TALUS_MI_DETECT_TEARS;
    if (every == 0) every = 1;
    ctx->ckpt.clear();
    ctx->ckpt_version = ~0ull;
    std::vector<uint8_t> torn;
    for (uint32_t t = 0; t < steps; ++t) {
      if (t % every == 0) {
        ctx->ckpt.emplace_back();
        ctx->engine->state_all_get(ctx->ckpt.back());
      }
      talus_mi_report r;
      ctx->engine->step(p, r, ctx->events, ctx->steps, torn);
      ++ctx->steps;
      report->bodies = r.bodies;
      report->nodes = r.nodes;
      report->elements = r.elements;
      report->substeps += r.substeps;
      report->cg_iterations += r.cg_iterations;
      report->cg_capped += r.cg_capped;
      report->solves += r.solves;
      report->refactors += r.refactors;
      report->cg_worst_relative = std::max(report->cg_worst_relative, r.cg_worst_relative);
      report->refinements += r.refinements;
      report->refine_unmet += r.refine_unmet;
      report->refine_worst_relative = std::max(report->refine_worst_relative, r.refine_worst_relative);
      report->elastic_energy = r.elastic_energy;
      report->kinetic_energy = r.kinetic_energy;
      report->yield_events += r.yield_events;
      report->preconditioner_applications += r.preconditioner_applications;
      report->device_seconds += r.device_seconds;
      if (r.non_finite) {
        report->non_finite = 1;
        return fail(TALUS_MI_ERR_STATE, "rollout became non-finite at step " + std::to_string(t));
      }
    }
    ctx->engine->state_all_get(ctx->ckpt_final);
    ctx->ckpt_steps = steps;
    ctx->ckpt_every = every;
    ctx->ckpt_params = p;
    ctx->ckpt_version = ctx->version;
    return TALUS_MI_OK;
  });
}

int talus_mi_rollout_backward(talus_mi_ctx *ctx, uint32_t body, const double *x_bar,
                              const double *v_bar, double tolerance, double *lambda_bar,
                              double *mu_bar, double *rest_bar, double *x0_bar, double *v0_bar,
                              double *a0_bar) {
  if (!ctx || !body_ok(ctx, body) || !x_bar) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  if (!(tolerance > 0.0)) return fail(TALUS_MI_ERR_ARGUMENT, "tolerance must be positive");
  return guarded(ctx, [&]() -> int {
    if (ctx->dirty || ctx->ckpt.empty() || ctx->ckpt_version != ctx->version)
      return fail(TALUS_MI_ERR_STATE, "no rollout of the current bodies to differentiate");
    const tmi::Body &b = ctx->bodies[body];
    for (uint32_t e = 0; e < b.tets; ++e)
      if ((b.flags[e] & (tmi::FLAG_PLASTIC | tmi::FLAG_DAMP)) || !(b.flags[e] & tmi::FLAG_SNH))
        return fail(TALUS_MI_ERR_ARGUMENT, "the adjoint supports undamped elastic neo-Hookean bodies only");
    // the replays overwrite the per-tet operator state: keep it
    ctx->engine->download(ctx->bodies);
    ctx->engine->rollout_backward(ctx->ckpt_params, ctx->ckpt, ctx->ckpt_final, ctx->ckpt_steps,
                                  ctx->ckpt_every, body, x_bar, v_bar, tolerance, lambda_bar,
                                  mu_bar, rest_bar, x0_bar, v0_bar, a0_bar);
    for (uint32_t q = 0; q < ctx->bodies.size(); ++q) {
      ctx->engine->rotations_set(q, ctx->bodies[q].R.data());
      ctx->engine->frames_set(q, ctx->bodies[q].W.data());
      ctx->engine->plastic_set(q, ctx->bodies[q].plastic.data());
    }
    return TALUS_MI_OK;
  });
}

// ---- strainstep hex lattices ----------------------------------------------

void talus_mi_lattice_default_params(talus_mi_lattice_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->dt = 1.0 / 60.0;
  p->gravity[1] = -9.81;
  p->softstep_substeps = 4;
}

namespace {

bool lattice_ok(talus_mi_ctx *ctx, uint32_t lattice) {
  return ctx->lattices && lattice < ctx->lattices->count();
}

// the lattice's screen, bound to its current mesh (host mirror downloaded)
tmi::HexScreen &screen_of(talus_mi_ctx *ctx, uint32_t lattice, tmi::LatticeBody *&body) {
  body = &ctx->lattices->body(lattice);
  if (ctx->screens.size() < ctx->lattices->count()) ctx->screens.resize(ctx->lattices->count());
  auto &slot = ctx->screens[lattice];
  if (!slot) slot = std::make_unique<tmi::HexScreen>(ctx->stream);
  slot->bind(*body);
  return *slot;
}

}  // namespace

int talus_mi_lattice_snh_terms(talus_mi_ctx *ctx, uint32_t lattice, const double *x, double *cell_energy,
                               double *node_gradient, double *node_hessian) {
  if (!ctx || !lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    tmi::LatticeBody *b = nullptr;
    tmi::HexScreen &sc = screen_of(ctx, lattice, b);
    sc.terms(x ? x : b->x.data(), cell_energy, node_gradient, node_hessian);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_energy(talus_mi_ctx *ctx, uint32_t lattice, const double *gravity, talus_mi_energy *out) {
  if (!ctx || !lattice_ok(ctx, lattice) || !gravity || !out) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    tmi::LatticeBody *b = nullptr;
    tmi::HexScreen &sc = screen_of(ctx, lattice, b);
    *out = sc.energy(b->x.data(), b->v.data(), gravity);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_rest_body(talus_mi_ctx *ctx, uint32_t lattice, talus_mi_rest_body *out) {
  if (!ctx || !lattice_ok(ctx, lattice) || !out) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    tmi::rest_body_of(ctx->lattices->body(lattice), *out);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_bake_field(talus_mi_ctx *ctx, uint32_t lattice, const talus_mi_sdf *rest,
                                const talus_mi_pose *pose, int carve_beyond_reach, talus_mi_sdf *current,
                                talus_mi_sample_box *changed) {
  if (!ctx || !lattice_ok(ctx, lattice) || !rest || !pose || !current || !changed || !rest->distances ||
      !current->distances)
    return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  for (int k = 0; k < 3; ++k)
    if (rest->resolution[k] < 2u) return fail(TALUS_MI_ERR_ARGUMENT, "a sampled field needs two samples per axis");
  return guarded(ctx, [&]() -> int {
    const tmi::LatticeBody &b = ctx->lattices->body(lattice);
    // model(pose) (pose.h:14-16): the rotation of the quaternion, then the
    // translation, in FP64 from the float pose
    const double w = pose->rotation[0], x = pose->rotation[1], y = pose->rotation[2], z = pose->rotation[3];
    const double model[12] = {1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y),
                              double(pose->position[0]),
                              2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x),
                              double(pose->position[1]),
                              2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y),
                              double(pose->position[2])};
    if (!tmi::bake_deformed_field(b, b.x.data(), model, *rest, carve_beyond_reach != 0, *current, *changed,
                                  ctx->stream))
      return fail(TALUS_MI_ERR_ARGUMENT, "the current field must share the rest field's lattice");
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_screen(talus_mi_ctx *ctx, uint32_t lattice, const uint32_t *held, uint32_t nheld,
                            const double *forces, double *displacement, double *dent, uint32_t *iterations) {
  if (!ctx || !lattice_ok(ctx, lattice) || !forces || !displacement || !dent || (nheld && !held))
    return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    tmi::LatticeBody *b = nullptr;
    tmi::HexScreen &sc = screen_of(ctx, lattice, b);
    for (uint32_t k = 0; k < nheld; ++k)
      if (held[k] >= b->nodes) return fail(TALUS_MI_ERR_ARGUMENT, "held node out of range");
    if (b->nodes == 0) return fail(TALUS_MI_ERR_ARGUMENT, "the lattice has no nodes");
    const uint32_t limit = std::max<uint32_t>(2000u, 30u * b->nodes);
    if (!sc.screen(*b, held, nheld, forces, displacement, dent, iterations, 1.0e-12, limit))
      return fail(TALUS_MI_ERR_STATE, "the screen solve did not converge");
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_add(talus_mi_ctx *ctx, const talus_mi_hex_mesh *mesh,
                         const talus_mi_material *materials, uint32_t lanes, const double *offset,
                         const double *velocity, uint32_t *lattice) {
  if (!mesh || !lattice) return fail(TALUS_MI_ERR_ARGUMENT, "null argument");
  return guarded(ctx, [&]() -> int {
    tmi::LatticeBody b;
    const int rc = tmi::make_lattice(b, mesh, materials, lanes, offset, velocity);
    if (rc != TALUS_MI_OK) return rc;
    if (!ctx->lattices) ctx->lattices = std::make_unique<tmi::LatticeSet>(ctx->stream);
    *lattice = ctx->lattices->add(std::move(b));
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_set_substeps(talus_mi_ctx *ctx, uint32_t lattice, uint32_t substeps) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    ctx->lattices->set_substeps(lattice, substeps);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_state_get(talus_mi_ctx *ctx, uint32_t lattice, double *x, double *v) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    ctx->lattices->get_state(lattice, x, v);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_state_set(talus_mi_ctx *ctx, uint32_t lattice, const double *x, const double *v) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    ctx->lattices->set_state(lattice, x, v);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_impulses_get(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *cells, double *impulses,
                                  uint32_t capacity, uint32_t *count) {
  if (!count) return fail(TALUS_MI_ERR_ARGUMENT, "null count");
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    const tmi::LatticeBody &b = ctx->lattices->body(lattice);
    const uint32_t n = uint32_t(b.elements.size());
    *count = n;
    if ((cells || impulses) && capacity < n) return fail(TALUS_MI_ERR_ARGUMENT, "capacity too small");
    for (uint32_t i = 0; i < n; ++i) {
      if (cells) cells[i] = b.elements[i].cell;
      if (impulses) {
        std::memcpy(impulses + 19 * size_t(i), b.elements[i].impulse, tmi::HEX_ROWS * sizeof(double));
        impulses[19 * size_t(i) + tmi::HEX_ROWS] = b.elements[i].vol_impulse;
      }
    }
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_plastic_get(talus_mi_ctx *ctx, uint32_t lattice, double *plastic, double *hardening,
                                 uint32_t *present) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    ctx->lattices->get_plastic(lattice, plastic, hardening, present);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_plastic_set(talus_mi_ctx *ctx, uint32_t lattice, const double *plastic,
                                 const double *hardening) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    ctx->lattices->set_plastic(lattice, plastic, hardening);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_colours_get(talus_mi_ctx *ctx, uint32_t lattice, uint8_t *node_colour,
                                 uint8_t *hex_colour) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    const tmi::LatticeBody &b = ctx->lattices->body(lattice);
    if (node_colour) std::memcpy(node_colour, b.node_colour.data(), b.nodes);
    if (hex_colour) std::memcpy(hex_colour, b.hex_colour.data(), b.hexes);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_topology_get(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *nodes, uint32_t *hex_nodes,
                                  float *node_mass, uint8_t *crack_faces) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    const tmi::LatticeBody &b = ctx->lattices->body(lattice);
    if (nodes) *nodes = b.nodes;
    if (hex_nodes) std::copy(b.hex.begin(), b.hex.end(), hex_nodes);
    if (node_mass) std::copy(b.mass.begin(), b.mass.end(), node_mass);
    if (crack_faces) {
      if (b.crack.empty())
        std::fill(crack_faces, crack_faces + b.hexes, uint8_t(0));
      else
        std::copy(b.crack.begin(), b.crack.end(), crack_faces);
    }
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_tear(talus_mi_ctx *ctx, uint32_t lattice, const uint32_t *cells, uint32_t count,
                          uint32_t *copies) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    if (count && !cells) return fail(TALUS_MI_ERR_ARGUMENT, "null cells");
    const uint32_t hexes = ctx->lattices->body(lattice).hexes;
    for (uint32_t i = 0; i < count; ++i)
      if (cells[i] >= hexes) return fail(TALUS_MI_ERR_ARGUMENT, "cell out of range");
    const uint32_t made = ctx->lattices->tear(lattice, cells, count);
    if (copies) *copies = made;
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_cell_stretch(talus_mi_ctx *ctx, uint32_t lattice, uint32_t cell, double *stretch,
                                  double *direction) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    if (!stretch || !direction) return fail(TALUS_MI_ERR_ARGUMENT, "null output");
    if (cell >= ctx->lattices->body(lattice).hexes) return fail(TALUS_MI_ERR_ARGUMENT, "cell out of range");
    ctx->lattices->cell_stretch(lattice, cell, stretch, direction);
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_lowest_get(talus_mi_ctx *ctx, uint32_t lattice, double *lowest) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    if (!lowest) return fail(TALUS_MI_ERR_ARGUMENT, "null output");
    *lowest = ctx->lattices->body(lattice).lowest;
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_surface_get(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *nodes, uint32_t capacity,
                                 uint32_t *count) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    if (!count) return fail(TALUS_MI_ERR_ARGUMENT, "null count");
    tmi::LatticeBody &b = ctx->lattices->body(lattice);
    b.ensure_surface();
    uint32_t n = 0;
    for (uint32_t i = 0; i < b.nodes; ++i) n += b.surface[i] ? 1u : 0u;
    *count = n;
    if (nodes) {
      if (capacity < n) return fail(TALUS_MI_ERR_ARGUMENT, "surface capacity too small");
      uint32_t k = 0;
      for (uint32_t i = 0; i < b.nodes; ++i)
        if (b.surface[i]) nodes[k++] = i;
    }
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_pieces(talus_mi_ctx *ctx, uint32_t lattice, uint32_t *pieces) {
  return guarded(ctx, [&]() -> int {
    if (!lattice_ok(ctx, lattice)) return fail(TALUS_MI_ERR_ARGUMENT, "no such lattice");
    if (!pieces) return fail(TALUS_MI_ERR_ARGUMENT, "null output");
    *pieces = ctx->lattices->body(lattice).pieces();
    return TALUS_MI_OK;
  });
}

int talus_mi_lattice_step(talus_mi_ctx *ctx, const talus_mi_lattice_params *params,
                          talus_mi_lattice_report *report) {
  if (!params) return fail(TALUS_MI_ERR_ARGUMENT, "null parameters");
  if (!(params->dt > 0.0) || !std::isfinite(params->dt)) return fail(TALUS_MI_ERR_ARGUMENT, "dt must be positive");
  return guarded(ctx, [&]() -> int {
    talus_mi_lattice_report r{};
    if (ctx->lattices) ctx->lattices->step(*params, r);
    if (report) *report = r;
    return TALUS_MI_OK;
  });
}

// ---------------------------------------------------------- contact queries

int talus_mi_field_add(talus_mi_ctx *ctx, const talus_mi_sdf *field, uint32_t *id) {
  if (!ctx || !field || !id) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    auto f = std::make_unique<tmi::DeviceField>();
    try {
      f->set(*field, ctx->stream);
    } catch (const std::invalid_argument &e) {
      return fail(TALUS_MI_ERR_ARGUMENT, e.what());
    }
    ctx->fields.push_back(std::move(f));
    *id = uint32_t(ctx->fields.size() - 1);
    return TALUS_MI_OK;
  });
}

int talus_mi_field_set(talus_mi_ctx *ctx, uint32_t id, const talus_mi_sdf *field) {
  if (!ctx || !field || id >= ctx->fields.size()) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  return guarded(ctx, [&]() -> int {
    try {
      ctx->fields[id]->set(*field, ctx->stream);
    } catch (const std::invalid_argument &e) {
      return fail(TALUS_MI_ERR_ARGUMENT, e.what());
    }
    return TALUS_MI_OK;
  });
}

int talus_mi_field_lipschitz(talus_mi_ctx *ctx, uint32_t id, float *lipschitz) {
  if (!ctx || id >= ctx->fields.size() || !lipschitz) return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  *lipschitz = ctx->fields[id]->lipschitz();
  return TALUS_MI_OK;
}

int talus_mi_find_contacts(talus_mi_ctx *ctx, uint32_t body, const talus_mi_shape *shapes, uint32_t count,
                           const talus_mi_contact_query *query, const float *thickness,
                           talus_mi_surface_contact *out, uint32_t capacity, uint32_t *found) {
  if (!ctx || !body_ok(ctx, body) || !query || !found || (count && !shapes) || (capacity && !out))
    return fail(TALUS_MI_ERR_ARGUMENT, "bad arguments");
  if (!(query->dt >= 0.0f) || !std::isfinite(query->dt)) return fail(TALUS_MI_ERR_ARGUMENT, "dt must be finite");
  std::vector<const tmi::DeviceField *> fields(count);
  for (uint32_t s = 0; s < count; ++s) {
    if (shapes[s].field >= ctx->fields.size()) return fail(TALUS_MI_ERR_ARGUMENT, "unknown field");
    fields[s] = ctx->fields[shapes[s].field].get();
  }
  return guarded(ctx, [&]() -> int {
    ensure(ctx);
    if (ctx->finders.size() != ctx->bodies.size()) {
      ctx->finders.resize(ctx->bodies.size());
      ctx->finder_version.resize(ctx->bodies.size(), ~0ull);
    }
    if (!ctx->finders[body] || ctx->finder_version[body] != ctx->version) {
      ctx->finders[body].reset();
      ctx->finders[body] = std::make_unique<tmi::ContactFinder>(ctx->bodies[body]);
      ctx->finder_version[body] = ctx->version;
    }
    const double *x = nullptr, *v = nullptr;
    hipStream_t s = nullptr;
    ctx->engine->device_state(body, &x, &v, &s);
    *found = ctx->finders[body]->find(x, v, s, shapes, fields.data(), count, *query, thickness, out, capacity);
    return TALUS_MI_OK;
  });
}
