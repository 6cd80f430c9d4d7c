// Montjoie's leapfrog driven one operator at a time through the C++ adapter
// (include/mjwave_adapter.hpp) on host vectors: the calls of
// ModifiedEquationIterator::SetInitialConditionS (order 2,
// Instationary/ModifiedEquation.cxx:99-124), of its Advance (:295-333) and
// the loop of RunSecondOrderScheme (:37-50), in the reference's order, with
// zero initial data.  Built and run by tests/test_gpu_boundary.py.
//
//   adapter_replay <input.bin> <states.bin>
// input.bin: int32 dim, order, nsteps, has_sigma; int64 n_nodes, n_elem;
//   double dt, f0; double xyz[N*dim]; int32 elem_nodes[E*nd];
//   uint8 dirichlet[N]; double rho[E], mu[E], sigma[E] (if has_sigma);
//   double src[dim].
// states.bin: U^n for n = 0..nsteps (N doubles each), the iterates that
//   GiveIterate / GiveFinalIterate receive.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mjwave_adapter.hpp"

template <class T>
static void rd(FILE *f, T *x, size_t n) {
  if (std::fread(x, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: adapter_replay input.bin states.bin\n");
    return 2;
  }
  FILE *in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t hdr[4];
  int64_t sz[2];
  double par[2];
  rd(in, hdr, 4);
  rd(in, sz, 2);
  rd(in, par, 2);
  const int dim = hdr[0], order = hdr[1], nsteps = hdr[2], has_sigma = hdr[3];
  const int64_t N = sz[0], E = sz[1];
  int nd = 1;
  for (int a = 0; a < dim; ++a) nd *= order + 1;
  const double dt = par[0], f0 = par[1];
  std::vector<double> xyz(N * dim), rho(E), mu(E), sigma(E, 0.0), src(dim);
  std::vector<int32_t> en(E * nd);
  std::vector<uint8_t> dir(N);
  rd(in, xyz.data(), xyz.size());
  rd(in, en.data(), en.size());
  rd(in, dir.data(), dir.size());
  rd(in, rho.data(), E);
  rd(in, mu.data(), E);
  if (has_sigma) rd(in, sigma.data(), E);
  rd(in, src.data(), dim);
  std::fclose(in);

  mjw::MeshView m{dim, order, N, E, xyz.data(), en.data(), dir.data()};
  mjw::HipWaveSystem sys(m, rho.data(), mu.data());
  if (has_sigma) sys.SetDamping(sigma);
  sys.SetDiracSources(src, dim);
  sys.SetRickerPulse(f0, 0.0, dt, nsteps);
  sys.SetTimeStep(dt);

  FILE *out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  const double t_begin = 0.0, dt2 = dt * dt;
  std::vector<double> U0(N, 0.0), U1(N), du0(N, 0.0), Ah(N, 0.0);

  // SetInitialConditionS (ModifiedEquation.cxx:99-124)
  double t = t_begin;
  U1 = U0;
  sys.SetDirichletCondition(t, 0, U0);
  sys.ApplyOperatorSh(1.0, t, du0, 0.0, U1);
  sys.ApplyOperatorKh(1.0, t, U0, -1.0, U1);
  sys.SolveOperatorDh(U1);
  for (int64_t i = 0; i < N; ++i) U1[i] = (0.5 * dt) * U1[i];
  for (int64_t i = 0; i < N; ++i) U1[i] += du0[i];
  for (int64_t i = 0; i < N; ++i) U1[i] = dt * U1[i];
  for (int64_t i = 0; i < N; ++i) U1[i] += U0[i];
  sys.SetDirichletCondition(t + dt, 0, U1);

  // RunSecondOrderScheme's loop (:37-50) around Advance (:295-333)
  for (int n = 0; n < nsteps; ++n) {
    t = t_begin + n * dt;
    std::fwrite(U0.data(), 8, N, out);  // GiveIterate(n, t, U0)
    sys.ApplyOperatorKh(1.0, t + dt, U1, 0.0, Ah);
    sys.AddScalarTimeSource(1.0, t + dt, 1, Ah);
    sys.ApplyOperatorDhMinusdtSh(-1.0, t + dt, U0, dt2, Ah);
    sys.ApplyOperatorDh(2.0, t + dt, U1, 1.0, Ah);
    U0 = U1;
    U1 = Ah;
    sys.SetDirichletConditionSource(t + 2 * dt, 0, U1);
    sys.SolveOperatorDhPlusdtSh(U1);
  }
  std::fwrite(U0.data(), 8, N, out);  // GiveFinalIterate
  std::fclose(out);
  return 0;
}
