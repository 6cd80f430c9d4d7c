// Compile check and usage example of include/mjwave_adapter.hpp (built by
// tests/test_abi.py with g++; run on a GPU box by tests/test_gpu_parity.py).
#include <cstdio>
#include <vector>

#include "mjwave_adapter.hpp"

int main() {
  const int dim = 2, nx = 8, order = 4;
  int64_t N = 0, E = 0;
  mjw::check(mjw_mesh_regular_size(dim, nx, order, &N, &E), nullptr, "size");
  std::vector<double> xyz(N * dim);
  std::vector<int32_t> en(E * (order + 1) * (order + 1));
  std::vector<uint8_t> dir(N);
  mjw::check(mjw_mesh_regular(dim, nx, order, 0.0, 0, xyz.data(), en.data(), dir.data()), nullptr,
             "mesh");
  mjw::MeshView m{dim, order, N, E, xyz.data(), en.data(), dir.data()};
  std::vector<double> rho(E, 1.0), mu(E, 1.0);
  mjw::HipWaveSystem sys(m, rho.data(), mu.data());
  sys.SetDiracSources({0.5, 0.5}, dim);
  sys.SetReceivers({0.8, 0.5, 0.3, 0.4}, dim);
  const double dt = 1e-3;
  const int nsteps = 200;
  sys.SetRickerPulse(10.0, 0.0, dt, nsteps);
  mjw::HipLeapFrogScheme lf(sys);
  lf.SetInitialConditionZero(0.0, dt);
  std::vector<double> traces(nsteps * sys.GetNbReceivers());
  lf.AdvanceN(nsteps, traces.data());
  for (int n = 100; n < 103; ++n) std::printf("%.15e %.15e\n", traces[2 * n], traces[2 * n + 1]);
  // the same run with the fourth-order scheme (OrderTimeScheme = 4 LEAP_FROG)
  sys.SetTemporalSource(MJW_TIME_RICKER, 10.0, 0.0, 0.0, dt, nsteps);
  lf.SetOrder(4);
  lf.SetInitialConditionZero(0.0, dt);
  lf.AdvanceN(nsteps, traces.data());
  for (int n = 100; n < 103; ++n) std::printf("%.15e %.15e\n", traces[2 * n], traces[2 * n + 1]);
  return 0;
}
