// mjwave_adapter.hpp -- header-only C++ adapter over the C ABI (mjwave.h),
// shaped like the two Montjoie interfaces the explicit driver talks to:
//
//  * the operator system VirtualOdeSystem<T> (Instationary/TimeSchemes.hxx:35-151,
//    implemented by VarInstationary_Base, VarInstationaryInline.cxx:174-242):
//    ApplyOperatorKh / ApplyOperatorDh / ApplyOperatorSh /
//    ApplyOperatorDhMinusdtSh / SolveOperatorDh / SolveOperatorDhPlusdtSh /
//    AddScalarTimeSource / SetDirichletCondition(Source), each on one host
//    vector (damping included: S = sigma J w after SetDamping);
//  * the leapfrog time scheme ModifiedEquationIterator<T>
//    (Instationary/ModifiedEquation.hxx:78-114, .cxx:99-124,295-345):
//    SetInitialConditionS / Advance / GetIterate / GetCFL / GetNumberOfIterates.
//
// Errors: the C ABI returns status codes; like the reference (cout + abort(),
// e.g. Hyperbolic/Acoustic/TimeAcoustic.cxx:243-249) this adapter prints the
// message and aborts, unless MJW_ADAPTER_THROW is defined (then it throws
// std::runtime_error).
//
// Vectors are std::vector<double> or any contiguous container exposing
// data() / size(); host memory stays owned by the caller.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "mjwave.h"

namespace mjw {

inline void check(int rc, const mjw_problem *p, const char *what) {
  if (rc == MJW_OK) return;
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " +
                    mjw_last_error(p);
#ifdef MJW_ADAPTER_THROW
  throw std::runtime_error(msg);
#else
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::abort();
#endif
}

// Mesh handed over by the caller: positions of all global GL nodes and the
// element -> node map in local order q = i + n j (+ n^2 k).
struct MeshView {
  int dim = 3, order = 1;
  int64_t n_nodes = 0, n_elem = 0;
  const double *xyz = nullptr;
  const int32_t *elem_nodes = nullptr;
  const uint8_t *dirichlet = nullptr;  // may be null (no Dirichlet node)
};

// The explicit second-order wave problem D u'' + A u = f(t) b on one device.
class HipWaveSystem {
 public:
  HipWaveSystem(const MeshView &m, const double *rho, const double *mu,
                const double *lambda = nullptr, int ncomp = 1, int nshots = 1,
                int precision = MJW_FP64, int device = 0)
      : ncomp_(ncomp), nshots_(nshots), ndof_(m.n_nodes * ncomp) {
    check(mjw_create(&p_, device, m.dim, m.order, ncomp, m.n_nodes, m.xyz, m.n_elem,
                     m.elem_nodes, rho, mu, lambda, m.dirichlet, nshots, precision),
          nullptr, "mjw_create");
  }
  ~HipWaveSystem() { mjw_destroy(p_); }
  HipWaveSystem(const HipWaveSystem &) = delete;
  HipWaveSystem &operator=(const HipWaveSystem &) = delete;

  mjw_problem *handle() const { return p_; }
  int64_t GetNbDof() const { return ndof_; }

  // Time step baked into D - dt/2 S and (D + dt/2 S)^{-1}, as Montjoie's
  // mass matrices take it at Init (TimeAcoustic.cxx:977-1004).
  void SetTimeStep(double dt) { dt_ = dt; }
  double GetTimeStep() const { return dt_; }
  // Damping sigma per element (MateriauDielec ... sigma): S_h = sigma J w
  void SetDamping(const std::vector<double> &sigma) {
    check(mjw_set_damping(p_, sigma.data()), p_, "SetDamping");
  }

  // Y = beta Y + alpha K U, K = -A  (ApplyOperatorKh)
  template <class V>
  void ApplyOperatorKh(double alpha, double /*t*/, const V &U, double beta, V &Y) {
    check(mjw_apply_kh(p_, alpha, U.data(), beta, Y.data()), p_, "ApplyOperatorKh");
  }
  // Y = beta Y + alpha D U  (ApplyOperatorDh)
  template <class V>
  void ApplyOperatorDh(double alpha, double /*t*/, const V &U, double beta, V &Y) {
    check(mjw_apply_dh(p_, alpha, U.data(), beta, Y.data()), p_, "ApplyOperatorDh");
  }
  // Y = beta Y + alpha S U  (ApplyOperatorSh)
  template <class V>
  void ApplyOperatorSh(double alpha, double /*t*/, const V &U, double beta, V &Y) {
    check(mjw_apply_sh(p_, alpha, U.data(), beta, Y.data()), p_, "ApplyOperatorSh");
  }
  // Y = beta Y + alpha (D - dt/2 S) U  (ApplyOperatorDhMinusdtSh)
  template <class V>
  void ApplyOperatorDhMinusdtSh(double alpha, double /*t*/, const V &U, double beta, V &Y) {
    check(mjw_apply_dh_minus_dt_sh(p_, dt_, alpha, U.data(), beta, Y.data()), p_,
          "ApplyOperatorDhMinusdtSh (SetTimeStep first)");
  }
  // Y = D^{-1} Y  (SolveOperatorDh)
  template <class V>
  void SolveOperatorDh(V &Y) {
    check(mjw_solve_dh(p_, Y.data()), p_, "SolveOperatorDh");
  }
  // Y = (D + dt/2 S)^{-1} Y  (SolveOperatorDhPlusdtSh)
  template <class V>
  void SolveOperatorDhPlusdtSh(V &Y) {
    check(mjw_solve_dh_plus_dt_sh(p_, dt_, Y.data()), p_,
          "SolveOperatorDhPlusdtSh (SetTimeStep first)");
  }
  // Y += alpha f^(nb_deriv)(t) F  (AddScalarTimeSource; shot 0)
  template <class V>
  void AddScalarTimeSource(double alpha, double t, int nb_deriv, V &Y) {
    check(mjw_add_scalar_time_source(p_, 0, alpha, t, nb_deriv, Y.data()), p_,
          "AddScalarTimeSource");
  }
  // homogeneous Dirichlet: zero the Dirichlet rows (SetDirichletCondition,
  // SetDirichletConditionSource)
  template <class V>
  void SetDirichletCondition(double /*t*/, int /*nb_deriv*/, V &Y) {
    check(mjw_set_dirichlet_condition(p_, Y.data()), p_, "SetDirichletCondition");
  }
  template <class V>
  void SetDirichletConditionSource(double /*t*/, int /*nb_deriv*/, V &Y) {
    check(mjw_set_dirichlet_condition(p_, Y.data()), p_, "SetDirichletConditionSource");
  }

  // Dirac sources with a Ricker pulse (SRC_DIRAC + TemporalSource = RICKER)
  void SetDiracSources(const std::vector<double> &pts, int dim,
                       const std::vector<double> &pol = {}, const std::vector<int32_t> &shot = {}) {
    const int nsrc = (int)(pts.size() / dim);
    check(mjw_set_sources(p_, nsrc, pts.data(), pol.empty() ? nullptr : pol.data(),
                          shot.empty() ? nullptr : shot.data()),
          p_, "SetDiracSources");
  }
  // TemporalSource = RICKER for both paths: the pulse table of the fused
  // device loop and the time function of AddScalarTimeSource (window
  // tolerance (final - initial time) * 1e-15, VarInstationary.cxx:993)
  void SetRickerPulse(double f0, double t_begin, double dt, int nsteps) {
    SetTemporalSource(MJW_TIME_RICKER, f0, 0.0, t_begin, dt, nsteps);
  }
  // any TemporalSource kind (MJW_TIME_*) with its value, for both
  void SetTemporalSource(int kind, double f0, double param, double t_begin, double dt,
                         int nsteps) {
    std::vector<double> tab(nsteps), tab2(nsteps), tab3(nsteps);
    check(mjw_temporal_source_table(kind, f0, param, t_begin, dt, nsteps, 1, dt, tab.data()),
          nullptr, "TemporalSource");
    check(mjw_set_pulse(p_, nsteps, tab.data()), p_, "SetTemporalSource");
    // f' and f'' for the fourth-order scheme (unused at order 2)
    check(mjw_temporal_source_table(kind, f0, param, t_begin, dt, nsteps, 2, dt, tab2.data()),
          nullptr, "TemporalSource");
    check(mjw_temporal_source_table(kind, f0, param, t_begin, dt, nsteps, 3, dt, tab3.data()),
          nullptr, "TemporalSource");
    check(mjw_set_pulse_derivatives(p_, nsteps, tab2.data(), tab3.data()), p_,
          "SetTemporalSource");
    check(mjw_set_temporal_source(p_, kind, f0, param, t_begin, (nsteps * dt) * 1e-15), p_,
          "SetTemporalSource");
  }
  void SetReceivers(const std::vector<double> &pts, int dim, const std::vector<int32_t> &comp = {}) {
    const int nrec = (int)(pts.size() / dim);
    check(mjw_set_receivers(p_, nrec, pts.data(), comp.empty() ? nullptr : comp.data()), p_,
          "SetReceivers");
    nrec_ = nrec;
  }
  int GetNbReceivers() const { return nrec_; }
  int GetNbShots() const { return nshots_; }

  // CFL: dt_max = 2 / sqrt(lambda_max(D^{-1} A))  (EvaluateCFL_Generic)
  double EvaluateCFL(int iters = 80) {
    double lam = 0.0;
    check(mjw_lambda_max(p_, iters, 1, &lam), p_, "EvaluateCFL");
    return 2.0 / std::sqrt(lam);
  }

 private:
  mjw_problem *p_ = nullptr;
  int ncomp_, nshots_;
  int64_t ndof_;
  int nrec_ = 0;
  double dt_ = 0.0;
};

// Order-2 leapfrog (ModifiedEquationIterator) whose iterates live on the
// device; Advance may fuse several steps (AdvanceN) and the receiver traces
// recorded at every step play the role of GiveIterate's seismograms.
class HipLeapFrogScheme {
 public:
  explicit HipLeapFrogScheme(HipWaveSystem &sys) : sys_(sys) {}

  // ModifiedEquationIterator::SetOrder (ModifiedEquation.cxx:71-79): 2 or 4;
  // order 4 also needs the source's f' and f'' (SetTemporalSource sets them)
  void SetOrder(int r) {
    check(mjw_set_time_order(sys_.handle(), r), sys_.handle(), "SetOrder");
    order_ = r;
  }

  // U^0 = u0, U^1 = U^0 + dt du0 + dt^2/2 D^{-1} K U^0 (+ Dirichlet)
  template <class V>
  void SetInitialConditionS(double t, double dt, const V &u0, const V &du0) {
    check(mjw_leapfrog_init(sys_.handle(), t, dt, u0.size() ? u0.data() : nullptr,
                            du0.size() ? du0.data() : nullptr),
          sys_.handle(), "SetInitialConditionS");
    dt_ = dt;
    sys_.SetTimeStep(dt);
  }
  void SetInitialConditionZero(double t, double dt) {
    check(mjw_leapfrog_init(sys_.handle(), t, dt, nullptr, nullptr), sys_.handle(),
          "SetInitialConditionS");
    dt_ = dt;
    sys_.SetTimeStep(dt);
  }
  // one step (Advance); traces of the iterate before the step appended
  void Advance(double /*t*/, int /*n*/) { AdvanceN(1, nullptr); }
  // n fused steps; traces: [n][nshots][nrec] or null
  void AdvanceN(int n, double *traces, bool graph = true) {
    check(mjw_leapfrog_advance(sys_.handle(), n, traces, graph ? 1 : 0), sys_.handle(),
          "Advance");
  }
  // U^n and U^{n+1} (GetIterate(0), GetIterate(1)); all shots
  void GetIterates(std::vector<double> &Un, std::vector<double> &Un1) {
    Un.resize(sys_.GetNbDof() * sys_.GetNbShots());
    Un1.resize(Un.size());
    check(mjw_get_state(sys_.handle(), Un.data(), Un1.data()), sys_.handle(), "GetIterate");
  }
  std::vector<double> Energy() {
    std::vector<double> e(sys_.GetNbShots());
    check(mjw_energy(sys_.handle(), e.data()), sys_.handle(), "Energy");
    return e;
  }
  // GetCFL, GetNumberOfIterates of ModifiedEquationIterator (:217-225, 264-270)
  double GetCFL() const { return order_ == 4 ? 2.0 * std::sqrt(3.0) : 2.0; }
  int GetNumberOfIterates() const { return order_ == 4 ? 3 : 2; }
  double GetTimeStep() const { return dt_; }

 private:
  HipWaveSystem &sys_;
  double dt_ = 0.0;
  int order_ = 2;
};

}  // namespace mjw
