// Host-side FP64 setup of the wave solver: 1-D Gauss-Lobatto rule and
// Lagrange basis, the Ricker pulse table, synthetic regular meshes, element
// colouring and point location.  Citations are to MONTJOIE/src of the
// reference.
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>

#include "mjw_internal.hpp"

namespace mjw {

// ---------------------------------------------------------------------------
// Gauss-Lobatto rule of [0,1] (Quadrature/GaussLobattoPoints.cxx:13-275,
// reached from Globatto::ConstructQuadrature, GaussLobatto.cxx:29-45):
// interior nodes are the zeros of P_r' (computed here by Newton's method on
// the Legendre recurrence in long double), weights 2/(r(r+1)P_r^2) on [-1,1],
// mapped to [0,1] and normalised to sum 1 (GaussLobattoPoints.cxx:262-270).
// ---------------------------------------------------------------------------
static void legendre(int r, long double t, long double &p, long double &dp, long double &d2p) {
  // P_k via recurrence; derivatives from (1-t^2) P_k' = k (P_{k-1} - t P_k)
  long double pm1 = 1.0L, pk = t;
  if (r == 0) { p = 1; dp = 0; d2p = 0; return; }
  for (int k = 2; k <= r; ++k) {
    long double pn = ((2.0L * k - 1.0L) * t * pk - (k - 1.0L) * pm1) / k;
    pm1 = pk;
    pk = pn;
  }
  p = pk;
  long double om = 1.0L - t * t;
  dp = r * (pm1 - t * pk) / om;
  d2p = (2.0L * t * dp - r * (r + 1.0L) * pk) / om;
}

void gl_rule(int order, double *x, double *w) {
  const int r = order;
  if (r == 1) { x[0] = 0; x[1] = 1; w[0] = w[1] = 0.5; return; }
  std::vector<long double> t(r + 1), wt(r + 1);
  t[0] = -1.0L;
  t[r] = 1.0L;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 1; k < r; ++k) {
    // Chebyshev-Lobatto guess, refined by Newton on P_r'
    long double z = -std::cos(pi * k / r);
    for (int it = 0; it < 200; ++it) {
      long double p, dp, d2p;
      legendre(r, z, p, dp, d2p);
      long double step = dp / d2p;
      z -= step;
      if (std::fabs(step) < 1e-20L) break;
    }
    t[k] = z;
  }
  std::sort(t.begin(), t.end());
  long double s = 0;
  for (int k = 0; k <= r; ++k) {
    long double p, dp, d2p;
    if (k == 0 || k == r) {
      p = (k == 0 && (r % 2)) ? -1.0L : 1.0L;  // P_r(+-1)
    } else {
      legendre(r, t[k], p, dp, d2p);
    }
    wt[k] = 2.0L / (r * (r + 1.0L) * p * p);
    s += wt[k];
  }
  for (int k = 0; k <= r; ++k) {
    // symmetric pair average keeps x_k + x_{r-k} = 1 and w_k = w_{r-k} exactly
    long double a = 0.5L * (1.0L + t[k]);
    long double b = 1.0L - 0.5L * (1.0L + t[r - k]);
    x[k] = (double)(0.5L * (a + b));
    w[k] = (double)(0.5L * (wt[k] + wt[r - k]) / s);
  }
  x[0] = 0.0;
  x[r] = 1.0;
  if (r == 2) { x[1] = 0.5; w[0] = w[2] = 1.0 / 6.0; w[1] = 2.0 / 3.0; }
}

// D[a*n+b] = phi_b'(x_a): barycentric form phi_b'(x_a) = (c_b/c_a)/(x_a - x_b)
// with c_k = 1/prod_{j!=k}(x_k-x_j), diagonal = -(row sum)
// (Globatto::ComputeGradPhi, Quadrature/GaussLobatto.cxx:175-192).
void gl_deriv(int order, const double *x, double *D) {
  const int n = order + 1;
  double c[64];
  for (int k = 0; k < n; ++k) {
    double p = 1.0;
    for (int j = 0; j < n; ++j)
      if (j != k) p *= (x[k] - x[j]);
    c[k] = 1.0 / p;
  }
  for (int a = 0; a < n; ++a) {
    double rs = 0.0;
    for (int b = 0; b < n; ++b) {
      if (b == a) continue;
      double v = c[b] / (c[a] * (x[a] - x[b]));
      D[a * n + b] = v;
      rs += v;
    }
    D[a * n + a] = (std::fabs(rs) <= 1e3 * 2.220446049250313e-16) ? 0.0 : -rs;
  }
}

// Lagrange basis on the nodes x (product form, GaussLobattoInline.cxx:83-117)
void lagrange(int order, const double *x, double t, double *phi, double *dphi) {
  const int n = order + 1;
  for (int i = 0; i < n; ++i) {
    double c = 1.0, p = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != i) { c *= (x[i] - x[k]); p *= (t - x[k]); }
    phi[i] = p / c;
    if (dphi) {
      double s = 0.0;
      for (int l = 0; l < n; ++l) {
        if (l == i) continue;
        double q = 1.0;
        for (int k = 0; k < n; ++k)
          if (k != i && k != l) q *= (t - x[k]);
        s += q;
      }
      dphi[i] = s / c;
    }
  }
}

// ---------------------------------------------------------------------------
// Time function of the sources: the primitive of the chosen pulse (the
// Evaluate of the VirtualTimeSource that GetNewTimeSource builds,
// Instationary/VarInstationary.cxx:409-481) interpolated by the degree-24 GL
// Lagrange polynomial on windows of length 0.5/f0 that move forward with t,
// derivatives by the GL differentiation matrix applied nb_deriv times
// (TimeSourceHyperbolic::Init / GenerateValues / EvaluateDerivative,
// Source/TimeSource.cxx:818-880).  nb_deriv = 1 is the pulse itself.
// ---------------------------------------------------------------------------
void GlTimeSource::init(double freq, double t_init, double eps_time, int kind_, double par_) {
  kind = kind_;
  par = par_;
  f0 = freq;
  eps = eps_time;
  t0 = t_init;
  t1 = t_init + 0.5 / freq;
  if (!ready) {
    double wg[kNP];
    gl_rule(kR, xg, wg);
    gl_deriv(kR, xg, Dg);
    for (int i = 0; i < kNP; ++i) {
      double p = 1.0;
      for (int k = 0; k < kNP; ++k)
        if (k != i) p *= (xg[i] - xg[k]);
      cinv[i] = 1.0 / p;
    }
    ready = true;
  }
  generate();
}

// The Evaluate of each pulse (Source/TimeSource.cxx): kind 1 RICKER, the
// Ricker primitive (:33-51); 2 DERIVATIVE_RICKER, the Ricker itself (:55-67,
// 121-124); 3 MODIFIED_RICKER centred at par (:152-160); 4 GAUSSIAN with a
// plateau of width par after one period (:198-213); 5 HARMONIC ending at par
// (:597-611); 6 SINUS_GAUSSIAN of parameter par centred half way to its
// cut-off (:667-679, VarInstationary.cxx:479-481).
double GlTimeSource::primitive(double t) const {
  const double pi = 3.14159265358979323846;
  switch (kind) {
    case 1:
    case 2: {
      const double T0 = 1.0 / f0;
      if (t <= -0.9 * T0 || t >= 2.9 * T0) return 0.0;
      const double a = f0 * t - 1.0;
      if (kind == 1) return -a * std::exp(-(a * a) * (pi * pi)) / f0;
      const double q = (a * pi) * (a * pi);
      return (2.0 * q - 1.0) * std::exp(-q);
    }
    case 3: {
      const double s = t - par, a = f0 * s;
      return -s * std::exp(-2.0 * ((a * pi) * (a * pi)));
    }
    case 4: {
      const double T0 = 1.0 / f0;
      double s;
      if (t < T0) s = t;
      else if (t >= T0 + par) s = t - par;
      else return 1.0;
      const double a = pi * (f0 * s - 1.0);
      return std::exp(-(a * a));
    }
    case 5: {
      if (t <= 0.0 || t >= par) return 0.0;
      const double a = f0 * f0 * t * t, b = f0 * f0 * (t - par) * (t - par);
      return std::exp(-1.0 / (pi * a)) * std::sin(2.0 * pi * f0 * t) * std::exp(-1.0 / (pi * b));
    }
    case 6: {
      const double toff = 0.5 * time_source_limit(6, f0, par), s = t - toff;
      return std::sin(2.0 * pi * f0 * t) * std::exp(-par * (s * s));
    }
    default:
      return 0.0;
  }
}

// tlimit_source of GetNewTimeSource (VarInstationary.cxx:418-481)
double time_source_limit(int kind, double f0, double par) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  if (!(f0 > 0.0) || !std::isfinite(f0) || !std::isfinite(par)) return nan;
  switch (kind) {
    case 1:
    case 2: return 2.5 / f0;
    case 3: return (2.5 + par) / (std::sqrt(2.0) * f0);
    case 4: return par >= 0.0 ? 2.5 / f0 + par : nan;
    case 5: return par;
    case 6: return par > 0.0 ? 2.0 * std::sqrt(std::log(1e6) / par) : nan;
    default: return nan;
  }
}

void GlTimeSource::generate() {
  for (int i = 0; i < kNP; ++i) val[0][i] = primitive((1.0 - xg[i]) * t0 + xg[i] * t1);
  const double inv = 1.0 / (t1 - t0);
  for (int d = 1; d < kNDer; ++d)
    for (int a = 0; a < kNP; ++a) {
      double s = 0.0;
      for (int k = 0; k < kNP; ++k) s += Dg[a * kNP + k] * val[d - 1][k];
      val[d][a] = s * inv;
    }
}

double GlTimeSource::eval(double t, int nb_deriv) {
  if (t > t1 + eps && t1 > t0) {
    while (t > t1 + eps) {
      const double len = t1 - t0;
      t0 = t1;
      t1 = t1 + len;
    }
    generate();
  }
  const double lam = (t - t0) / (t1 - t0);
  double acc = 0.0;
  for (int i = 0; i < kNP; ++i) {
    double p = 1.0;
    for (int k = 0; k < kNP; ++k)
      if (k != i) p *= (lam - xg[k]);
    acc += (p * cinv[i]) * val[nb_deriv][i];
  }
  return acc;
}

// Pulse table of the leapfrog: the value at step n is the first derivative at
// t_begin + n*dt + dt (ModifiedEquation.cxx:316), window tolerance
// (final - initial time) * 1e-15 (VarInstationary.cxx:993), zero from the
// cut-off 2.5/f0 on (VarInstationary.cxx:416-420, 2394).
void ricker_pulse_table(double f0, double t_begin, double dt, int nsteps, double *out) {
  source_table(1, f0, 0.0, t_begin, dt, nsteps, 1, dt, out);
}

void source_table(int kind, double f0, double par, double t_begin, double dt, int nsteps,
                  int nb_deriv, double offset, double *out) {
  GlTimeSource src;
  src.init(f0, t_begin, (nsteps * dt) * 1e-15, kind, par);
  const double tcut = time_source_limit(kind, f0, par);
  for (int n = 0; n < nsteps; ++n) {
    const double t = (t_begin + n * dt) + offset;
    out[n] = t >= tcut ? 0.0 : src.eval(t, nb_deriv);
  }
}

// ---------------------------------------------------------------------------
// Synthetic regular meshes of [0,1]^dim (REGULAR meshes, Mesh/Mesh3D.cxx:
// 8406-8486: vertex i at i*h, last at 1) with seeded interior-vertex
// perturbation.  Nodes of each Q_r element are multilinear images of the GL
// nodes; the global node grid is lexicographic.
// ---------------------------------------------------------------------------
static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline double unif(uint64_t seed, uint64_t idx) {
  const uint64_t h = mix64(mix64(seed) ^ (idx * 0xD1B54A32D192ED03ULL));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

void mesh_regular_size(int dim, int nx, int order, int64_t *N, int64_t *E) {
  const int64_t g = (int64_t)nx * order + 1;
  *N = dim == 2 ? g * g : g * g * g;
  *E = dim == 2 ? (int64_t)nx * nx : (int64_t)nx * nx * nx;
}

void mesh_regular(int dim, int nx, int order, double pert, uint64_t seed, double *xyz,
                  int32_t *elem_nodes, uint8_t *dirichlet) {
  const int n = order + 1;
  const int nv1 = nx + 1;
  const int64_t nv = dim == 2 ? (int64_t)nv1 * nv1 : (int64_t)nv1 * nv1 * nv1;
  std::vector<double> vert(nv * dim);
  const double h = 1.0 / nx, scale = pert * h / std::sqrt((double)dim);
  for (int64_t v = 0; v < nv; ++v) {
    int idx[3] = {(int)(v % nv1), (int)((v / nv1) % nv1), (int)(v / ((int64_t)nv1 * nv1))};
    bool inner = true;
    for (int a = 0; a < dim; ++a) inner = inner && idx[a] > 0 && idx[a] < nx;
    for (int a = 0; a < dim; ++a) {
      double c = idx[a] == nx ? 1.0 : idx[a] * (1.0 / nx);
      if (inner && pert != 0.0) c += (2.0 * unif(seed, 3 * (uint64_t)v + a) - 1.0) * scale;
      vert[v * dim + a] = c;
    }
  }
  double xg[kMaxN + 1], wg[kMaxN + 1];
  gl_rule(order, xg, wg);
  const int64_t g1 = (int64_t)nx * order + 1;
  int64_t N, E;
  mesh_regular_size(dim, nx, order, &N, &E);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < N; ++g) {
    const int64_t I[3] = {g % g1, (g / g1) % g1, dim == 3 ? g / (g1 * g1) : 0};
    int el[3] = {0, 0, 0};
    double xi[3] = {0, 0, 0};
    bool bnd = false;
    for (int a = 0; a < dim; ++a) {
      el[a] = (int)std::min<int64_t>(I[a] / order, nx - 1);
      xi[a] = xg[I[a] - (int64_t)el[a] * order];
      bnd = bnd || I[a] == 0 || I[a] == g1 - 1;
    }
    double acc[3] = {0, 0, 0};
    for (int c = 0; c < (1 << dim); ++c) {
      double wgt = 1.0;
      int64_t vid = 0, stride = 1;
      for (int a = 0; a < dim; ++a) {
        const int bit = (c >> a) & 1;
        wgt *= bit ? xi[a] : 1.0 - xi[a];
        vid += (el[a] + bit) * stride;
        stride *= nv1;
      }
      for (int a = 0; a < dim; ++a) acc[a] += wgt * vert[vid * dim + a];
    }
    for (int a = 0; a < dim; ++a) xyz[g * dim + a] = acc[a];
    dirichlet[g] = bnd ? 1 : 0;
  }
  const int nd = dim == 2 ? n * n : n * n * n;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    const int64_t ex = e % nx, ey = (e / nx) % nx, ez = dim == 3 ? e / ((int64_t)nx * nx) : 0;
    for (int q = 0; q < nd; ++q) {
      const int i = q % n, j = (q / n) % n, k = q / (n * n);
      elem_nodes[e * nd + q] =
          (int32_t)((ex * order + i) + g1 * ((ey * order + j) + (dim == 3 ? g1 * (ez * order + k) : 0)));
    }
  }
}

// ---------------------------------------------------------------------------
// Element colouring: greedy first fit with a per-node bitmask of used colours.
// Montjoie's element loop is a serial scatter (ProdMatVectScalarH1.cxx:715-717,
// NumberMap.cxx:745-777); colours make it race-free without atomics.
// ---------------------------------------------------------------------------
int color_elements(int64_t E, int nd, int64_t N, const int32_t *elem_nodes,
                   std::vector<int> &color) {
  std::vector<uint64_t> used(N, 0);
  color.assign(E, 0);
  int nc = 0;
  for (int64_t e = 0; e < E; ++e) {
    uint64_t m = 0;
    const int32_t *en = elem_nodes + e * nd;
    for (int q = 0; q < nd; ++q) m |= used[en[q]];
    if (~m == 0) return -1;
    const int c = __builtin_ctzll(~m);
    color[e] = c;
    nc = std::max(nc, c + 1);
    for (int q = 0; q < nd; ++q) used[en[q]] |= (1ULL << c);
  }
  return nc;
}

// ---------------------------------------------------------------------------
// Diagonal operators on host vectors [ncomp][N] (the lumped mass, damping and
// Dirichlet entries of the operator-level boundary): per node coefficient
// a[g] + cc * c[g] (a null: 0, c null: a alone), in parallel over the nodes.
// ---------------------------------------------------------------------------
void diag_axpby(int64_t N, int ncomp, const double *a, const double *c, double cc, double alpha,
                const double *U, double beta, double *Y) {
  for (int m = 0; m < ncomp; ++m) {
    const double *u = U + (int64_t)m * N;
    double *y = Y + (int64_t)m * N;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < N; ++g) {
      const double coef = a ? (c ? a[g] + cc * c[g] : a[g]) : 0.0;
      const double d = alpha * (coef * u[g]);
      y[g] = beta == 0.0 ? d : beta * y[g] + d;
    }
  }
}

void diag_solve(int64_t N, int ncomp, const double *a, const double *c, double cc, double *Y) {
  for (int m = 0; m < ncomp; ++m) {
    double *y = Y + (int64_t)m * N;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < N; ++g) y[g] *= 1.0 / (c ? a[g] + cc * c[g] : a[g]);
  }
}

void diag_mask(int64_t N, int ncomp, const uint8_t *dir, double *Y) {
  for (int m = 0; m < ncomp; ++m) {
    double *y = Y + (int64_t)m * N;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < N; ++g)
      if (dir[g]) y[g] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Straight-sided hexahedra: for a non-curved element Montjoie evaluates DF of
// the transformation built from the element's vertices at the quadrature
// points (ComputeLocalMassMatrix, Harmonic/VarGeometryProblem.cxx:671-690,
// GetVerticesElement + DFjElem), i.e. of the trilinear map of the 8 vertices.
// When every node is that map's image of the GL grid the kernels rebuild DF
// from 21 coefficients per element instead of reading 6 stored values per
// node.
// ---------------------------------------------------------------------------
bool trilinear_coefficients(int order, int64_t E, const int32_t *elem_nodes, const double *xyz,
                            const int64_t *perm, const double *mu_sorted, double tol,
                            std::vector<double> &coef) {
  const int n = order + 1, nd = n * n * n, r = order;
  double xg[kMaxN], wg[kMaxN];
  gl_rule(order, xg, wg);
  coef.assign((size_t)E * kCoef, 0.0);
  // elements are independent: fitted in parallel, the verdict shared
  int bad = 0;
#pragma omp parallel for schedule(static)
  for (int64_t es = 0; es < E; ++es) {
    int failed;
#pragma omp atomic read
    failed = bad;
    if (failed) continue;
    const int32_t *en = elem_nodes + perm[es] * nd;
    auto X = [&](int i, int j, int k, int d) {
      return xyz[(int64_t)en[i + n * j + n * n * k] * 3 + d];
    };
    double a[8][3];
    double diam2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double x000 = X(0, 0, 0, d), x100 = X(r, 0, 0, d), x010 = X(0, r, 0, d),
                   x110 = X(r, r, 0, d), x001 = X(0, 0, r, d), x101 = X(r, 0, r, d),
                   x011 = X(0, r, r, d), x111 = X(r, r, r, d);
      a[0][d] = x000;
      a[1][d] = x100 - x000;
      a[2][d] = x010 - x000;
      a[3][d] = x001 - x000;
      a[4][d] = x110 - x100 - x010 + x000;
      a[5][d] = x011 - x010 - x001 + x000;
      a[6][d] = x101 - x100 - x001 + x000;
      a[7][d] = x111 - x110 - x101 - x011 + x100 + x010 + x001 - x000;
      diam2 += (x111 - x000) * (x111 - x000);
    }
    const double lim = tol * std::sqrt(diam2);
    bool fits = true;
    for (int k = 0; k < n && fits; ++k)
      for (int j = 0; j < n && fits; ++j)
        for (int i = 0; i < n && fits; ++i) {
          const double u = xg[i], v = xg[j], w = xg[k];
          for (int d = 0; d < 3; ++d) {
            const double x = a[0][d] + a[1][d] * u + a[2][d] * v + a[3][d] * w +
                             a[4][d] * u * v + a[5][d] * v * w + a[6][d] * u * w +
                             a[7][d] * u * v * w;
            if (!(std::fabs(x - X(i, j, k, d)) <= lim)) fits = false;
          }
        }
    if (!fits) {
#pragma omp atomic write
      bad = 1;
      continue;
    }
    double *c = &coef[(size_t)es * kCoef];
    for (int m = 1; m < 8; ++m)
      for (int d = 0; d < 3; ++d) c[(m - 1) * 3 + d] = a[m][d];
    c[21] = mu_sorted[es];
  }
  return !bad;
}

// ---------------------------------------------------------------------------
// Point location in an isoparametric element (replaces FjInverseProblem::Solve,
// FiniteElement/PointsReference.cxx:277-360, and the projection of points found
// slightly outside, DefineSourceElliptic.cxx:2539-2543).
// ---------------------------------------------------------------------------
// gradient weights at xi: g[a*nd + q] = d phi_q / d x_a = sum_b DF^{-1}[b][a]
// d phi_q / d xi_b, DF = sum_q x_q (grad_xi phi_q)^T the Jacobian of the
// element's nodal map (exact for the multilinear map of straight elements)
void tensor_grad(int dim, int order, const double *xq, const double *xi, double *g) {
  const int n = order + 1;
  const int nd = dim == 2 ? n * n : n * n * n;
  double xg[kMaxN + 1], wg[kMaxN + 1], ph[3][kMaxN + 1], dph[3][kMaxN + 1];
  gl_rule(order, xg, wg);
  for (int a = 0; a < dim; ++a) lagrange(order, xg, xi[a], ph[a], dph[a]);
  std::vector<double> gr((size_t)nd * 3);
  double J[9] = {0};
  for (int q = 0; q < nd; ++q) {
    const int ii[3] = {q % n, (q / n) % n, q / (n * n)};
    for (int b = 0; b < dim; ++b) {
      double t = dph[b][ii[b]];
      for (int c = 0; c < dim; ++c)
        if (c != b) t *= ph[c][ii[c]];
      gr[(size_t)q * 3 + b] = t;
    }
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b) J[a * dim + b] += gr[(size_t)q * 3 + b] * xq[q * dim + a];
  }
  double inv[9];
  if (dim == 2) {
    const double det = J[0] * J[3] - J[1] * J[2];
    inv[0] = J[3] / det; inv[1] = -J[1] / det; inv[2] = -J[2] / det; inv[3] = J[0] / det;
  } else {
    const double c0 = J[4] * J[8] - J[5] * J[7], c1 = J[5] * J[6] - J[3] * J[8],
                 c2 = J[3] * J[7] - J[4] * J[6];
    const double det = J[0] * c0 + J[1] * c1 + J[2] * c2;
    const double adj[9] = {c0, J[2] * J[7] - J[1] * J[8], J[1] * J[5] - J[2] * J[4],
                           c1, J[0] * J[8] - J[2] * J[6], J[2] * J[3] - J[0] * J[5],
                           c2, J[1] * J[6] - J[0] * J[7], J[0] * J[4] - J[1] * J[3]};
    for (int k = 0; k < 9; ++k) inv[k] = adj[k] / det;
  }
  for (int q = 0; q < nd; ++q)
    for (int a = 0; a < dim; ++a) {
      double v = 0.0;
      for (int b = 0; b < dim; ++b) v += inv[b * dim + a] * gr[(size_t)q * 3 + b];
      g[(size_t)a * nd + q] = v;
    }
}

void tensor_basis(int dim, int order, const double *xi, double *phi) {
  const int n = order + 1;
  double xg[kMaxN + 1], wg[kMaxN + 1], p[3][kMaxN + 1];
  gl_rule(order, xg, wg);
  for (int a = 0; a < dim; ++a) lagrange(order, xg, xi[a], p[a], nullptr);
  const int nd = dim == 2 ? n * n : n * n * n;
  for (int q = 0; q < nd; ++q) {
    const int i = q % n, j = (q / n) % n, k = q / (n * n);
    phi[q] = p[0][i] * p[1][j] * (dim == 3 ? p[2][k] : 1.0);
  }
}

bool locate_point(int dim, int order, const double *xq, const double *p, double *xi) {
  const int n = order + 1;
  const int nd = dim == 2 ? n * n : n * n * n;
  double xg[kMaxN + 1], wg[kMaxN + 1];
  gl_rule(order, xg, wg);
  for (int a = 0; a < dim; ++a) xi[a] = 0.5;
  double ph[3][kMaxN + 1], dph[3][kMaxN + 1];
  for (int it = 0; it < 60; ++it) {
    for (int a = 0; a < dim; ++a) lagrange(order, xg, xi[a], ph[a], dph[a]);
    double F[3] = {0, 0, 0}, J[9] = {0};
    for (int q = 0; q < nd; ++q) {
      const int ii[3] = {q % n, (q / n) % n, q / (n * n)};
      double v = 1.0, d[3];
      for (int b = 0; b < dim; ++b) v *= ph[b][ii[b]];
      for (int b = 0; b < dim; ++b) {
        double t = dph[b][ii[b]];
        for (int c = 0; c < dim; ++c)
          if (c != b) t *= ph[c][ii[c]];
        d[b] = t;
      }
      for (int a = 0; a < dim; ++a) {
        F[a] += v * xq[q * dim + a];
        for (int b = 0; b < dim; ++b) J[a * dim + b] += d[b] * xq[q * dim + a];
      }
    }
    for (int a = 0; a < dim; ++a) F[a] -= p[a];
    double delta[3];
    if (dim == 2) {
      const double det = J[0] * J[3] - J[1] * J[2];
      delta[0] = (J[3] * F[0] - J[1] * F[1]) / det;
      delta[1] = (-J[2] * F[0] + J[0] * F[1]) / det;
    } else {
      const double c0 = J[4] * J[8] - J[5] * J[7], c1 = J[5] * J[6] - J[3] * J[8],
                   c2 = J[3] * J[7] - J[4] * J[6];
      const double det = J[0] * c0 + J[1] * c1 + J[2] * c2;
      const double inv[9] = {c0, J[2] * J[7] - J[1] * J[8], J[1] * J[5] - J[2] * J[4],
                             c1, J[0] * J[8] - J[2] * J[6], J[2] * J[3] - J[0] * J[5],
                             c2, J[1] * J[6] - J[0] * J[7], J[0] * J[4] - J[1] * J[3]};
      for (int a = 0; a < 3; ++a)
        delta[a] = (inv[a * 3] * F[0] + inv[a * 3 + 1] * F[1] + inv[a * 3 + 2] * F[2]) / det;
    }
    double nrm = 0;
    for (int a = 0; a < dim; ++a) {
      xi[a] -= delta[a];
      nrm += delta[a] * delta[a];
      if (!std::isfinite(xi[a])) return false;
      xi[a] = std::max(-1.0, std::min(2.0, xi[a]));
    }
    if (nrm < 1e-30) break;
  }
  const double tol = 1e-10;
  for (int a = 0; a < dim; ++a)
    if (xi[a] < -tol || xi[a] > 1.0 + tol) return false;
  for (int a = 0; a < dim; ++a) xi[a] = std::max(0.0, std::min(1.0, xi[a]));
  return true;
}

}  // namespace mjw
