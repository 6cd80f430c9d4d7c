// mjwave_montjoie.hpp -- the Montjoie-side binding: an implementation of
// Montjoie's operator interface VirtualOdeSystem<Real_wp>
// (Instationary/TimeSchemes.hxx:35-151) whose operators run in libmjwave.so.
// Any Montjoie time scheme that uses only these operators -- the order-2
// leapfrog ModifiedEquationIterator (Instationary/ModifiedEquation.cxx:
// 99-124,295-345) driven by RunSecondOrderScheme (:16-53) -- then steps the
// problem held by an mjw::HipWaveSystem without further changes.
//
// Include it inside a Montjoie build, after Montjoie's headers
// (MontjoieFlag.hxx + Instationary/MontjoieTime.hxx or any header that pulls
// them in); vectors are Montjoie's Seldon Vector<Real_wp>, Real_wp = double.
// Operators the explicit path does not use keep the base class's behaviour.
#pragma once

#include <functional>

#include "mjwave_adapter.hpp"

namespace Montjoie {

class HipOdeSystem : public VirtualOdeSystem<Real_wp> {
 public:
  // `on_iterate(n, t, U)` receives the iterates that RunSecondOrderScheme
  // hands to GiveIterate / GiveFinalIterate (seismograms, snapshots)
  explicit HipOdeSystem(mjw::HipWaveSystem &sys,
                        std::function<void(int, Real_wp, const VectReal_wp &)> on_iterate = {})
      : sys_(sys), p_(sys.handle()), on_iterate_(std::move(on_iterate)) {}

  // RunSecondOrderScheme announces dt before SetInitialConditionS
  // (ModifiedEquation.cxx:32): the dt of D - dt/2 S and (D + dt/2 S)^{-1}
  void GiveNumberIterations(const Real_wp &dt, int) override { sys_.SetTimeStep(dt); }
  void GiveIterate(int n, const Real_wp &t, VectReal_wp &U) override {
    if (on_iterate_) on_iterate_(n, t, U);
  }
  void GiveFinalIterate(int n, const Real_wp &t, VectReal_wp &U) override {
    if (on_iterate_) on_iterate_(n, t, U);
  }

  // Y = beta Y + alpha K U, K = -A  (VarInstationaryInline.cxx:174-179)
  void ApplyOperatorKh(const Real_wp &alpha, const Real_wp &, const VectReal_wp &U,
                       const Real_wp &beta, VectReal_wp &Y) override {
    mjw::check(mjw_apply_kh(p_, alpha, U.GetData(), beta, Y.GetData()), p_, "ApplyOperatorKh");
  }
  // lumped mass operators (ContinuousUnsteadyMassMatrix.cxx:1309-1420)
  void ApplyOperatorDh(const Real_wp &alpha, const Real_wp &, const VectReal_wp &U,
                       const Real_wp &beta, VectReal_wp &Y) override {
    mjw::check(mjw_apply_dh(p_, alpha, U.GetData(), beta, Y.GetData()), p_, "ApplyOperatorDh");
  }
  void ApplyOperatorSh(const Real_wp &alpha, const Real_wp &, const VectReal_wp &U,
                       const Real_wp &beta, VectReal_wp &Y) override {
    mjw::check(mjw_apply_sh(p_, alpha, U.GetData(), beta, Y.GetData()), p_, "ApplyOperatorSh");
  }
  void ApplyOperatorDhMinusdtSh(const Real_wp &alpha, const Real_wp &, const VectReal_wp &U,
                                const Real_wp &beta, VectReal_wp &Y) override {
    mjw::check(mjw_apply_dh_minus_dt_sh(p_, sys_.GetTimeStep(), alpha, U.GetData(), beta,
                                        Y.GetData()),
               p_, "ApplyOperatorDhMinusdtSh");
  }
  void SolveOperatorDh(VectReal_wp &Y) override {
    mjw::check(mjw_solve_dh(p_, Y.GetData()), p_, "SolveOperatorDh");
  }
  void SolveOperatorDhPlusdtSh(VectReal_wp &Y) override {
    mjw::check(mjw_solve_dh_plus_dt_sh(p_, sys_.GetTimeStep(), Y.GetData()), p_,
               "SolveOperatorDhPlusdtSh");
  }
  // Y += alpha (D + dt/2 S)^{-1} X
  void SolveOperatorDhPlusdtSh(const Real_wp &alpha, const VectReal_wp &X,
                               VectReal_wp &Y) override {
    VectReal_wp Z(X);
    SolveOperatorDhPlusdtSh(Z);
    for (int i = 0; i < Y.GetM(); i++) Y(i) += alpha * Z(i);
  }
  // homogeneous Dirichlet rows set to 0 (VarInstationary.cxx:702-725,2466-2505)
  void SetDirichletCondition(const Real_wp &, int, VectReal_wp &Y, Real_wp) override {
    mjw::check(mjw_set_dirichlet_condition(p_, Y.GetData()), p_, "SetDirichletCondition");
  }
  void SetDirichletConditionSource(const Real_wp &, int, VectReal_wp &Y, Real_wp) override {
    mjw::check(mjw_set_dirichlet_condition(p_, Y.GetData()), p_, "SetDirichletConditionSource");
  }
  // Y += alpha f^(n)(t) F  (VarInstationary.cxx:2391-2420,2509-2513)
  void AddScalarTimeSource(const Real_wp &alpha, const Real_wp &t, int n,
                           VectReal_wp &Y) override {
    mjw::check(mjw_add_scalar_time_source(p_, 0, alpha, t, n, Y.GetData()), p_,
               "AddScalarTimeSource");
  }

 private:
  mjw::HipWaveSystem &sys_;
  mjw_problem *p_;
  std::function<void(int, Real_wp, const VectReal_wp &)> on_iterate_;
};

}  // namespace Montjoie
