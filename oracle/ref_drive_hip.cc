// Montjoie's own order-2 leapfrog (ModifiedEquationIterator, driven by
// RunSecondOrderScheme, Instationary/ModifiedEquation.cxx:16-53,99-124,
// 295-345), compiled from the reference sources where they lie, stepping a
// problem whose operators run in libmjwave.so through the binding
// include/mjwave_montjoie.hpp (HipOdeSystem : VirtualOdeSystem<Real_wp>).
// Test infrastructure only (see Makefile, target "drive"): it checks that
// the library is a drop-in for the operator interface of Montjoie's time
// schemes.
//
//   mjref_drive_hip.x <input.bin> <states.bin>
// input.bin: int32 dim, order, nsteps, has_sigma; int64 n_nodes, n_elem;
//   double dt, f0; double xyz[N*dim]; int32 elem_nodes[E*nd];
//   uint8 dirichlet[N]; double rho[E], mu[E], sigma[E] (if has_sigma);
//   double src[dim].
// states.bin: the iterates GiveIterate / GiveFinalIterate receive, n = 0..nsteps.
#include "MontjoieFlag.hxx"
#include "Montjoie.hxx"
using namespace Montjoie;
#include "mjwave_montjoie.hpp"

template <class T>
static void rd(FILE *f, T *x, size_t n) {
  if (fread(x, sizeof(T), n, f) != n) { cout << "short input" << endl; exit(2); }
}

int main(int argc, char **argv)
{
  if (argc != 3) { cout << "usage: mjref_drive_hip.x input.bin states.bin" << endl; return 2; }
  FILE *in = fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t hdr[4]; int64_t sz[2]; double par[2];
  rd(in, hdr, 4); rd(in, sz, 2); rd(in, par, 2);
  const int dim = hdr[0], order = hdr[1], nsteps = hdr[2], has_sigma = hdr[3];
  const int64_t N = sz[0], E = sz[1];
  int nd = 1;
  for (int a = 0; a < dim; a++) nd *= order + 1;
  const double dt = par[0], f0 = par[1];
  std::vector<double> xyz(N * dim), rho(E), mu(E), sigma(E, 0.0), src(dim);
  std::vector<int32_t> en(E * nd);
  std::vector<uint8_t> dir(N);
  rd(in, xyz.data(), xyz.size()); rd(in, en.data(), en.size()); rd(in, dir.data(), dir.size());
  rd(in, rho.data(), E); rd(in, mu.data(), E);
  if (has_sigma) rd(in, sigma.data(), E);
  rd(in, src.data(), dim);
  fclose(in);

  mjw::MeshView m;
  m.dim = dim; m.order = order; m.n_nodes = N; m.n_elem = E;
  m.xyz = xyz.data(); m.elem_nodes = en.data(); m.dirichlet = dir.data();
  mjw::HipWaveSystem hip(m, rho.data(), mu.data());
  if (has_sigma) hip.SetDamping(sigma);
  hip.SetDiracSources(src, dim);
  hip.SetRickerPulse(f0, 0.0, dt, nsteps);

  FILE *out = fopen(argv[2], "wb");
  if (!out) return 2;
  HipOdeSystem sys(hip, [out](int, Real_wp, const VectReal_wp& U)
                   { fwrite(U.GetData(), sizeof(Real_wp), U.GetM(), out); });
  ModifiedEquationIterator<Real_wp> scheme;
  scheme.SetOrder(2);
  VectReal_wp U0(N), dU0(N);
  U0.Zero(); dU0.Zero();
  // t_end as a data file's TimeInterval gives nb_iter = ceil(t_end/dt) = nsteps
  RunSecondOrderScheme(Real_wp(0), Real_wp((nsteps - 0.5) * dt), Real_wp(dt), U0, dU0, sys,
                       scheme);
  fclose(out);
  return 0;
}
