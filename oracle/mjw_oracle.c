/*
 * mjw_oracle.c -- CPU FP64 restatement of Montjoie's explicit spectral-element
 * leapfrog path (acoustic 2-D/3-D and isotropic elastic 2-D/3-D).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and by nothing else.  The product library
 * (montjoie_amd/csrc) never links, loads or calls it.
 *
 * Every routine names the reference file:line it restates (paths relative to
 * MONTJOIE/src of the reference).  Arithmetic is plain FP64, element loops are
 * OpenMP-parallel inside one colour class (elements of one colour share no
 * node), so results are deterministic for a fixed thread count.
 *
 * Conventions shared with the product library (they are data formats, so they
 * have to be alike):
 *   - 1-D nodes x[0..r] are the Gauss-Lobatto points on [0,1], weights sum to 1.
 *   - D[a*n+b] = phi_b'(x_a)   (derivative of basis b at node a).
 *   - element-local node q = i + n*j (+ n*n*k), i along xi_1.
 *   - geometric factor per element-local node: the symmetric tensor
 *     G = J w DF^{-1} mu DF^{-T}, stored (xx,xy,yy) in 2-D and
 *     (xx,xy,xz,yy,yz,zz) in 3-D;  lumped mass  M = rho J w.
 *   - stiffness A: (A u)_i = sum_e  sum_q  grad(phi_i) . G_q grad(u)  so that
 *     Montjoie's K_h = -A.
 *
 * Parity status of each routine is listed in the header block of
 * oracle/README-oracle.txt (pinned / unpinned and by what).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXN 10   /* largest element nodes per direction (order <= 9) */
#define OR_MAXN1D 32 /* largest 1-D rule (the GL-24 pulse interpolator uses 25) */

/* ------------------------------------------------------------------------ */
/* 1-D Gauss-Lobatto rule                                                   */
/* ------------------------------------------------------------------------ */

/* Legendre P_m(t) and P_m'(t), P_m''(t) on [-1,1] by the three-term
 * recurrence (long double). */
static void legendre_pdd(int m, long double t, long double *p, long double *dp,
                         long double *ddp) {
  long double p0 = 1.0L, p1 = t, dp0 = 0.0L, dp1 = 1.0L, ddp0 = 0.0L, ddp1 = 0.0L;
  if (m == 0) { *p = 1.0L; *dp = 0.0L; *ddp = 0.0L; return; }
  for (int k = 2; k <= m; k++) {
    long double p2 = ((2 * k - 1) * t * p1 - (k - 1) * p0) / k;
    long double dp2 = dp0 + (2 * k - 1) * p1;          /* P_k' = P_{k-2}' + (2k-1)P_{k-1} */
    long double ddp2 = ddp0 + (2 * k - 1) * dp1;
    p0 = p1; p1 = p2; dp0 = dp1; dp1 = dp2; ddp0 = ddp1; ddp1 = ddp2;
  }
  *p = p1; *dp = dp1; *ddp = ddp1;
}

/*
 * Gauss-Lobatto points and weights on [0,1] for order r (r+1 points), weights
 * normalised to sum 1.  Restates Quadrature/GaussLobattoPoints.cxx:13-275
 * (ComputeGaussFormulas with gauss_lobatto=true, reached from
 * Quadrature/GaussJacobi.cxx:178-181 and GaussLobatto.cxx:29-45): the interior
 * points are the roots of P_r', the weights are 2/(r(r+1) P_r(t)^2) on
 * [-1,1], then mapped by x = (1+t)/2 and divided by their sum
 * (GaussLobattoPoints.cxx:262-270).  The reference finds the roots by a
 * Newton iteration in theta on a Fourier form of P_r; this restatement uses a
 * Newton iteration in t on the recurrence, in long double -- the roots are the
 * same numbers.  Closed forms for r = 1, 2 as in GaussLobattoPoints.cxx:62-85.
 */
int or_gl(int r, double *x, double *w) {
  int n = r + 1;
  if (r < 1 || n > OR_MAXN1D) return -1;
  if (r == 1) { x[0] = 0.0; x[1] = 1.0; w[0] = 0.5; w[1] = 0.5; return 0; }
  if (r == 2) {
    x[0] = 0.0; x[1] = 0.5; x[2] = 1.0;
    w[0] = 1.0 / 6.0; w[1] = 2.0 / 3.0; w[2] = w[0];
    return 0;
  }
  long double t[OR_MAXN1D], ww[OR_MAXN1D];
  t[0] = -1.0L; t[r] = 1.0L;
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 1; k < r; k++) {
    /* Chebyshev-Gauss-Lobatto initial guess */
    long double z = -cosl(pi * k / r);
    for (int it = 0; it < 100; it++) {
      long double p, dp, ddp;
      legendre_pdd(r, z, &p, &dp, &ddp);
      long double dz = dp / ddp;
      z -= dz;
      if (fabsl(dz) <= 1e-19L * (1.0L + fabsl(z))) break;
    }
    t[k] = z;
  }
  long double sum = 0.0L;
  for (int k = 0; k <= r; k++) {
    long double p, dp, ddp;
    legendre_pdd(r, t[k], &p, &dp, &ddp);
    ww[k] = 2.0L / ((long double)r * (r + 1) * p * p);
    sum += ww[k];
  }
  /* symmetrise exactly (the rule is symmetric about 1/2) */
  for (int k = 0; k <= r; k++) {
    long double xa = 0.5L + 0.5L * t[k];
    long double xb = 0.5L + 0.5L * t[r - k];
    x[k] = (double)(0.5L * (xa + (1.0L - xb)));
    w[k] = (double)(0.5L * (ww[k] + ww[r - k]) / sum);
  }
  x[0] = 0.0; x[r] = 1.0;
  return 0;
}

/* cte_phi(i) = 1 / prod_{j != i} (x_i - x_j)  (GaussLobatto.cxx:96-108) */
static void cte_phi(int r, const double *x, double *c) {
  for (int i = 0; i <= r; i++) {
    double p = 1.0;
    for (int j = 0; j <= r; j++)
      if (j != i) p *= x[i] - x[j];
    c[i] = 1.0 / p;
  }
}

/*
 * Derivative matrix D[a*n+b] = phi_b'(x_a).  Restates
 * Globatto::ComputeGradPhi (GaussLobatto.cxx:175-192): off-diagonal
 * val_grad(j,i) = cte_phi(j) / (cte_phi(i) (x_i - x_j)) = phi_j'(x_i),
 * diagonal = minus the row sum, zeroed below eps (1e3 machine eps, the value
 * passed from HexahedronGauss.cxx:292).
 */
void or_gl_deriv(int r, const double *x, double *D) {
  int n = r + 1;
  double c[OR_MAXN1D];
  cte_phi(r, x, c);
  for (int i = 0; i < n; i++) {
    double sum = 0.0;
    for (int j = 0; j < n; j++)
      if (j != i) {
        double v = c[j] / (c[i] * (x[i] - x[j]));
        D[i * n + j] = v;
        sum += v;
      }
    double d = -sum;
    if (fabs(d) <= 1e3 * 2.220446049250313e-16) d = 0.0;
    D[i * n + i] = d;
  }
}

/* phi_i(t) = cte_phi(i) prod_{k != i} (t - x_k)  (GaussLobattoInline.cxx:83-90) */
double or_phi(int r, const double *x, int i, double t) {
  double c[OR_MAXN1D];
  cte_phi(r, x, c);
  double p = 1.0;
  for (int k = 0; k <= r; k++)
    if (k != i) p *= t - x[k];
  return p * c[i];
}

/* ------------------------------------------------------------------------ */
/* Synthetic meshes                                                         */
/* ------------------------------------------------------------------------ */

/* splitmix64: seeded, counter-based uniform doubles in [0,1). */
static uint64_t smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double or_uniform(uint64_t seed, uint64_t idx) {
  uint64_t h = smix(smix(seed) ^ (idx * 0xD1B54A32D192ED03ULL));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

/* vertex coordinate along one axis: REGULAR mesh of [0,1] with nx+1 points
 * (Mesh/Mesh3D.cxx:8453-8456,8530-8546: pos = xmin + i*step, last = xmax). */
static double reg_coord(int i, int nx) {
  if (i == nx) return 1.0;
  return 0.0 + i * ((1.0 - 0.0) / nx);
}

/*
 * Vertices of a regular mesh of [0,1]^dim with nx elements per direction,
 * interior vertices displaced by pert*h*(2u-1)/sqrt(dim) per coordinate
 * (|displacement| <= pert*h), u from splitmix64(seed, 3*vertex + coord).
 * vert[v*dim + a], v = a0 + (nx+1)*(a1 + (nx+1)*a2).
 */
void or_mesh_vertices(int dim, int nx, double pert, uint64_t seed, double *vert) {
  int nv1 = nx + 1;
  long nv = (dim == 2) ? (long)nv1 * nv1 : (long)nv1 * nv1 * nv1;
  double h = 1.0 / nx, sd = 1.0 / sqrt((double)dim);
  for (long v = 0; v < nv; v++) {
    int ia[3] = {(int)(v % nv1), (int)((v / nv1) % nv1), (int)(v / ((long)nv1 * nv1))};
    int interior = 1;
    for (int a = 0; a < dim; a++)
      if (ia[a] == 0 || ia[a] == nx) interior = 0;
    for (int a = 0; a < dim; a++) {
      double c = reg_coord(ia[a], nx);
      if (interior && pert != 0.0)
        c += pert * h * (2.0 * or_uniform(seed, 3 * (uint64_t)v + a) - 1.0) * sd;
      vert[v * dim + a] = c;
    }
  }
}

/* multilinear shape functions of the unit square/cube; corner c = a+2b(+4c) */
static void multilinear(int dim, const double *xi, double *N, double *dN) {
  int nc = 1 << dim;
  for (int c = 0; c < nc; c++) {
    double f[3], g[3];
    for (int a = 0; a < dim; a++) {
      int bit = (c >> a) & 1;
      f[a] = bit ? xi[a] : 1.0 - xi[a];
      g[a] = bit ? 1.0 : -1.0;
    }
    double v = 1.0;
    for (int a = 0; a < dim; a++) v *= f[a];
    N[c] = v;
    for (int b = 0; b < dim; b++) {
      double d = g[b];
      for (int a = 0; a < dim; a++)
        if (a != b) d *= f[a];
      dN[c * dim + b] = d;
    }
  }
}

/* corner vertex ids of structured element e (ex fastest) */
static void elem_corners(int dim, int nx, long e, long *cv) {
  int nv1 = nx + 1;
  int ex = (int)(e % nx), ey = (int)((e / nx) % nx), ez = (dim == 3) ? (int)(e / ((long)nx * nx)) : 0;
  for (int c = 0; c < (1 << dim); c++) {
    int a = ex + (c & 1), b = ey + ((c >> 1) & 1), z = ez + ((c >> 2) & 1);
    cv[c] = a + (long)nv1 * (b + (long)nv1 * z);
  }
}

/*
 * Structured Q_r mesh: global GL node grid (nx*r+1)^dim numbered
 * lexicographically, element e = ex + nx*(ey + nx*ez), local node
 * q = i + n*(j + n*k) -> global (ex*r+i, ey*r+j, ez*r+k).  Node positions are
 * the multilinear image of the GL points of one canonical element
 * (min(I/r, nx-1)), the way Montjoie places nodes of straight Q1-geometry
 * elements (Fb.FjElem, Harmonic/VarGeometryProblem.cxx:693).  dirichlet[g]=1
 * on the boundary of [0,1]^dim.
 */
int or_mesh_nodes(int dim, int nx, int r, const double *vert, double *xyz,
                  int *elem_nodes, unsigned char *dirichlet) {
  int n = r + 1;
  double x[OR_MAXN], w[OR_MAXN];
  if (or_gl(r, x, w)) return -1;
  long ng1 = (long)nx * r + 1;
  long N = (dim == 2) ? ng1 * ng1 : ng1 * ng1 * ng1;
  long E = (dim == 2) ? (long)nx * nx : (long)nx * nx * nx;
  int nd = (dim == 2) ? n * n : n * n * n;
#pragma omp parallel for schedule(static)
  for (long g = 0; g < N; g++) {
    long I[3] = {g % ng1, (g / ng1) % ng1, (dim == 3) ? g / (ng1 * ng1) : 0};
    int el[3] = {0, 0, 0}, loc[3] = {0, 0, 0};
    for (int a = 0; a < dim; a++) {
      el[a] = (int)(I[a] / r);
      if (el[a] > nx - 1) el[a] = nx - 1;
      loc[a] = (int)(I[a] - (long)el[a] * r);
    }
    long e = el[0] + (long)nx * (el[1] + (dim == 3 ? (long)nx * el[2] : 0));
    long cv[8];
    elem_corners(dim, nx, e, cv);
    double xi[3], N8[8], dN[24];
    for (int a = 0; a < dim; a++) xi[a] = x[loc[a]];
    multilinear(dim, xi, N8, dN);
    for (int a = 0; a < dim; a++) {
      double s = 0.0;
      for (int c = 0; c < (1 << dim); c++) s += N8[c] * vert[cv[c] * dim + a];
      xyz[g * dim + a] = s;
    }
    int bnd = 0;
    for (int a = 0; a < dim; a++)
      if (I[a] == 0 || I[a] == ng1 - 1) bnd = 1;
    dirichlet[g] = (unsigned char)bnd;
  }
#pragma omp parallel for schedule(static)
  for (long e = 0; e < E; e++) {
    long ex = e % nx, ey = (e / nx) % nx, ez = (dim == 3) ? e / ((long)nx * nx) : 0;
    for (int q = 0; q < nd; q++) {
      int i = q % n, j = (q / n) % n, k = (dim == 3) ? q / (n * n) : 0;
      long gI = ex * r + i, gJ = ey * r + j, gK = ez * r + k;
      elem_nodes[e * nd + q] = (int)(gI + ng1 * (gJ + (dim == 3 ? ng1 * gK : 0)));
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Geometry, stiffness coefficients, lumped mass                            */
/* ------------------------------------------------------------------------ */

static double det_inv(int dim, const double *A, double *Ai) {
  if (dim == 2) {
    double d = A[0] * A[3] - A[1] * A[2];
    Ai[0] = A[3] / d; Ai[1] = -A[1] / d; Ai[2] = -A[2] / d; Ai[3] = A[0] / d;
    return d;
  }
  double c00 = A[4] * A[8] - A[5] * A[7];
  double c01 = A[5] * A[6] - A[3] * A[8];
  double c02 = A[3] * A[7] - A[4] * A[6];
  double d = A[0] * c00 + A[1] * c01 + A[2] * c02;
  Ai[0] = c00 / d; Ai[1] = (A[2] * A[7] - A[1] * A[8]) / d; Ai[2] = (A[1] * A[5] - A[2] * A[4]) / d;
  Ai[3] = c01 / d; Ai[4] = (A[0] * A[8] - A[2] * A[6]) / d; Ai[5] = (A[2] * A[3] - A[0] * A[5]) / d;
  Ai[6] = c02 / d; Ai[7] = (A[1] * A[6] - A[0] * A[7]) / d; Ai[8] = (A[0] * A[4] - A[1] * A[3]) / d;
  return d;
}

/*
 * Per element-local node: DF from the multilinear map of the element's
 * corner vertices (Fb.DFjElem, VarGeometryProblem.cxx:693-695), J = det DF,
 * DF^{-1} (VarGeometryProblem.cxx:750-756), then the stiffness tensor
 * G = J w DF^{-1} mu DF^{-T} and the mass coefficient rho J w
 * (Elliptic/Helmholtz/VarHelmholtz.cxx:88-117, second-order branch).
 * Also the elastic factors: Jw and DF^{-1} (row-major) for
 * Computation/ProdMatVectH1.cxx:760-820.
 * Inputs: structured mesh (dim, nx, r, vert), mu and rho per element
 * (nodal = 0) or per element-local node (nodal = 1, [E*nd]: media varying
 * inside the element, evaluated at the quadrature points as
 * VarGeometryProblem.cxx:1395-1560 stores them for VarHelmholtz.cxx:73-75).
 * Outputs (any may be NULL): G[E*nd*ncomp], Mloc[E*nd], jw[E*nd],
 * dfinv[E*nd*dim*dim].  Returns min J (negative = degenerate element).
 */
double or_geometry(int dim, int nx, int r, const double *vert, const double *mu,
                   const double *rho, int nodal, double *G, double *Mloc, double *jw,
                   double *dfinv) {
  int n = r + 1;
  double x[OR_MAXN], w[OR_MAXN];
  if (or_gl(r, x, w)) return -1.0;
  long E = (dim == 2) ? (long)nx * nx : (long)nx * nx * nx;
  int nd = (dim == 2) ? n * n : n * n * n;
  int nc = (dim == 2) ? 3 : 6;
  double jmin = 1e300;
#pragma omp parallel for schedule(static) reduction(min : jmin)
  for (long e = 0; e < E; e++) {
    long cv[8];
    elem_corners(dim, nx, e, cv);
    for (int q = 0; q < nd; q++) {
      int loc[3] = {q % n, (q / n) % n, q / (n * n)};
      double xi[3], N8[8], dN[24], DF[9], Di[9];
      double wq = 1.0;
      for (int a = 0; a < dim; a++) { xi[a] = x[loc[a]]; wq *= w[loc[a]]; }
      multilinear(dim, xi, N8, dN);
      for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) {
          double s = 0.0;
          for (int c = 0; c < (1 << dim); c++) s += dN[c * dim + b] * vert[cv[c] * dim + a];
          DF[a * dim + b] = s; /* dX_a / dxi_b */
        }
      double J = det_inv(dim, DF, Di);
      if (J < jmin) jmin = J;
      double Jw = J * wq;
      long o = e * nd + q;
      const double rho_q = nodal ? rho[o] : rho[e], mu_q = nodal ? mu[o] : mu[e];
      if (Mloc) Mloc[o] = rho_q * Jw;
      if (jw) jw[o] = Jw;
      if (dfinv)
        for (int a = 0; a < dim * dim; a++) dfinv[o * dim * dim + a] = Di[a];
      if (G) {
        /* G_ab = Jw mu sum_c Di[a][c] Di[b][c] */
        double gm[9] = {0};
        for (int a = 0; a < dim; a++)
          for (int b = 0; b < dim; b++) {
            double s = 0.0;
            for (int c = 0; c < dim; c++) s += Di[a * dim + c] * Di[b * dim + c];
            gm[a * dim + b] = Jw * mu_q * s;
          }
        if (dim == 2) {
          G[o * 3 + 0] = gm[0]; G[o * 3 + 1] = gm[1]; G[o * 3 + 2] = gm[3];
        } else {
          G[o * 6 + 0] = gm[0]; G[o * 6 + 1] = gm[1]; G[o * 6 + 2] = gm[2];
          G[o * 6 + 3] = gm[4]; G[o * 6 + 4] = gm[5]; G[o * 6 + 5] = gm[8];
        }
      }
    }
  }
  (void)nc;
  return jmin;
}

/*
 * Assembled lumped mass: M_g = sum_e rho J w over the elements holding g;
 * Dirichlet rows set to 1 (Hyperbolic/Acoustic/TimeAcoustic.cxx:931-937,
 * SetCoefficientDirichlet(1) + AddMatrixWithBC), and its inverse
 * (TimeAcoustic.cxx:955-965).
 */
void or_mass(long E, int nd, long N, const int *elem_nodes, const double *Mloc,
             const unsigned char *dirichlet, double *M, double *Minv) {
  for (long g = 0; g < N; g++) M[g] = 0.0;
  for (long e = 0; e < E; e++)
    for (int q = 0; q < nd; q++) M[elem_nodes[e * nd + q]] += Mloc[e * nd + q];
  for (long g = 0; g < N; g++) {
    if (dirichlet && dirichlet[g]) M[g] = 1.0;
    if (Minv) Minv[g] = 1.0 / M[g];
  }
}

/* ------------------------------------------------------------------------ */
/* Element colouring (no two elements of one colour share a node)           */
/* ------------------------------------------------------------------------ */

/* Greedy first-fit colouring over elements in index order.  Returns the
 * number of colours; color[e] in [0, ncolor). */
int or_color(long E, int nd, long N, const int *elem_nodes, int *color) {
  /* node -> bitmask of colours already used by elements touching it */
  uint64_t *used = (uint64_t *)calloc((size_t)N, sizeof(uint64_t));
  int ncolor = 0;
  for (long e = 0; e < E; e++) {
    uint64_t m = 0;
    for (int q = 0; q < nd; q++) m |= used[elem_nodes[e * nd + q]];
    int c = 0;
    while (c < 64 && ((m >> c) & 1)) c++;
    if (c == 64) { free(used); return -1; }
    color[e] = c;
    if (c + 1 > ncolor) ncolor = c + 1;
    for (int q = 0; q < nd; q++) used[elem_nodes[e * nd + q]] |= (1ULL << c);
  }
  free(used);
  return ncolor;
}

/* ------------------------------------------------------------------------ */
/* Matrix-free stiffness (scalar)                                           */
/* ------------------------------------------------------------------------ */

/*
 * Local product out = A_e u_e for one element, restating
 * MltAdd_ElementH1_Scalar (Computation/ProdMatVectScalarH1.cxx:256-668) on
 * the Lobatto path: Ch = I (FiniteElement/Hexahedron/HexahedronLobatto.cxx:
 * 199-203), gradient = Rh^T u (HexahedronLobatto.cxx:257-281 with Rh from
 * HexahedronGauss.cxx:243-364: entry phi_i'(xi_l)), per-node tensor
 * (ProdMatVectScalarH1.cxx:472-495), divergence = Rh v
 * (HexahedronLobatto.cxx:217-251).
 */
static void local_apply(int dim, int n, const double *D, const double *Gq,
                        const double *ue, double *out) {
  int nd = (dim == 2) ? n * n : n * n * n;
  double v[3 * OR_MAXN * OR_MAXN * OR_MAXN];
  for (int q = 0; q < nd; q++) {
    int i = q % n, j = (q / n) % n, k = q / (n * n);
    double g[3] = {0.0, 0.0, 0.0};
    for (int l = 0; l < n; l++) {
      g[0] += D[i * n + l] * ue[l + n * j + n * n * k];
      g[1] += D[j * n + l] * ue[i + n * l + n * n * k];
      if (dim == 3) g[2] += D[k * n + l] * ue[i + n * j + n * n * l];
    }
    if (dim == 2) {
      const double *G = Gq + 3 * q;
      v[2 * q + 0] = G[0] * g[0] + G[1] * g[1];
      v[2 * q + 1] = G[1] * g[0] + G[2] * g[1];
    } else {
      const double *G = Gq + 6 * q;
      v[3 * q + 0] = G[0] * g[0] + G[1] * g[1] + G[2] * g[2];
      v[3 * q + 1] = G[1] * g[0] + G[3] * g[1] + G[4] * g[2];
      v[3 * q + 2] = G[2] * g[0] + G[4] * g[1] + G[5] * g[2];
    }
  }
  for (int q = 0; q < nd; q++) {
    int i = q % n, j = (q / n) % n, k = q / (n * n);
    double s = 0.0;
    for (int l = 0; l < n; l++) {
      if (dim == 2) {
        s += D[l * n + i] * v[2 * (l + n * j) + 0];
        s += D[l * n + j] * v[2 * (i + n * l) + 1];
      } else {
        s += D[l * n + i] * v[3 * (l + n * j + n * n * k) + 0];
        s += D[l * n + j] * v[3 * (i + n * l + n * n * k) + 1];
        s += D[l * n + k] * v[3 * (i + n * j + n * n * l) + 2];
      }
    }
    out[q] = s;
  }
}

/*
 * y = beta*y + alpha * A u over all elements (MltAddScalarH1,
 * ProdMatVectScalarH1.cxx:679-727; C2.Fill(0)/Mlt(beta) at :693-696, element
 * loop :715-717, scatter-add via AddLocalUnknownVector, Mesh/NumberMap.cxx:
 * 745-777).  Elements are visited colour by colour so that the scatter is
 * race-free under OpenMP.  color may be NULL (serial loop).
 */
void or_stiffness(int dim, int r, long E, long N, const int *elem_nodes, const double *G,
                  const int *color, int ncolor, double alpha, const double *u, double beta,
                  double *y) {
  int n = r + 1;
  int nd = (dim == 2) ? n * n : n * n * n;
  int nc = (dim == 2) ? 3 : 6;
  double x[OR_MAXN], w[OR_MAXN], D[OR_MAXN * OR_MAXN];
  or_gl(r, x, w);
  or_gl_deriv(r, x, D);
  if (beta == 0.0)
    for (long g = 0; g < N; g++) y[g] = 0.0;
  else if (beta != 1.0)
    for (long g = 0; g < N; g++) y[g] *= beta;
  if (!color) {
    double ue[OR_MAXN * OR_MAXN * OR_MAXN], out[OR_MAXN * OR_MAXN * OR_MAXN];
    for (long e = 0; e < E; e++) {
      for (int q = 0; q < nd; q++) ue[q] = u[elem_nodes[e * nd + q]];
      local_apply(dim, n, D, G + e * nd * nc, ue, out);
      for (int q = 0; q < nd; q++) y[elem_nodes[e * nd + q]] += alpha * out[q];
    }
    return;
  }
  for (int c = 0; c < ncolor; c++) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < E; e++) {
      if (color[e] != c) continue;
      double ue[OR_MAXN * OR_MAXN * OR_MAXN], out[OR_MAXN * OR_MAXN * OR_MAXN];
      for (int q = 0; q < nd; q++) ue[q] = u[elem_nodes[e * nd + q]];
      local_apply(dim, n, D, G + e * nd * nc, ue, out);
      for (int q = 0; q < nd; q++) y[elem_nodes[e * nd + q]] += alpha * out[q];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Matrix-free stiffness (isotropic elastic, d components)                  */
/* ------------------------------------------------------------------------ */

/*
 * Local elastic product, restating MltAdd_ElementH1
 * (Computation/ProdMatVectH1.cxx:674-944) for the second-order continuous
 * case: three reference gradients (Rh^T, :745-751), physical gradient
 * dU = DF^{-T} grad (:762-777, there scaled by J and unscaled at :803),
 * sigma = C : dU with isotropic C (lambda tr(e) I + 2 mu e, e = sym(dU);
 * ApplyTensorStiffness, Elliptic/Elastic/VarElastic.cxx:22-30),
 * v = J w DF^{-1} sigma (:806-820), three divergences (:839-856).
 * Unknown layout: component-blocked global vector u[m*N + g]
 * (Mesh/NumberMap.cxx:693).
 */
static void local_apply_elastic(int n, const double *D, const double *jw, const double *dfi,
                                double lam, double mu, const double *ue, double *out) {
  int nd = n * n * n;
  double *v = (double *)malloc(sizeof(double) * 9 * nd);
  for (int q = 0; q < nd; q++) {
    int i = q % n, j = (q / n) % n, k = q / (n * n);
    double gr[3][3]; /* gr[m][b] = d u_m / d xi_b */
    for (int m = 0; m < 3; m++) {
      const double *um = ue + m * nd;
      double g0 = 0, g1 = 0, g2 = 0;
      for (int l = 0; l < n; l++) {
        g0 += D[i * n + l] * um[l + n * j + n * n * k];
        g1 += D[j * n + l] * um[i + n * l + n * n * k];
        g2 += D[k * n + l] * um[i + n * j + n * n * l];
      }
      gr[m][0] = g0; gr[m][1] = g1; gr[m][2] = g2;
    }
    const double *Di = dfi + 9 * q; /* Di[a*3+b] = (DF^{-1})_{ab} */
    double du[3][3];                /* du[m][a] = d u_m / d x_a = sum_b Di[b][a] gr[m][b] */
    for (int m = 0; m < 3; m++)
      for (int a = 0; a < 3; a++)
        du[m][a] = Di[0 * 3 + a] * gr[m][0] + Di[1 * 3 + a] * gr[m][1] + Di[2 * 3 + a] * gr[m][2];
    double tr = du[0][0] + du[1][1] + du[2][2];
    double sg[3][3];
    for (int m = 0; m < 3; m++)
      for (int a = 0; a < 3; a++) sg[m][a] = mu * (du[m][a] + du[a][m]) + (m == a ? lam * tr : 0.0);
    double s = jw[q];
    for (int m = 0; m < 3; m++)
      for (int b = 0; b < 3; b++)
        v[9 * q + 3 * m + b] =
            s * (Di[b * 3 + 0] * sg[m][0] + Di[b * 3 + 1] * sg[m][1] + Di[b * 3 + 2] * sg[m][2]);
  }
  for (int q = 0; q < nd; q++) {
    int i = q % n, j = (q / n) % n, k = q / (n * n);
    for (int m = 0; m < 3; m++) {
      double s = 0.0;
      for (int l = 0; l < n; l++) {
        s += D[l * n + i] * v[9 * (l + n * j + n * n * k) + 3 * m + 0];
        s += D[l * n + j] * v[9 * (i + n * l + n * n * k) + 3 * m + 1];
        s += D[l * n + k] * v[9 * (i + n * j + n * n * l) + 3 * m + 2];
      }
      out[m * nd + q] = s;
    }
  }
  free(v);
}

/*
 * The same product on quadrilaterals (the 2-D instance of MltAdd_ElementH1,
 * Computation/ProdMatVectH1.cxx:674-944, with ApplyTensorStiffness of
 * Elliptic/Elastic/VarElastic.cxx:22-30 in 2-D): two components, reference
 * gradients along xi_1 and xi_2, DF^{-1} 2x2 (row-major), plane strain
 * sigma = lambda tr(e) I + 2 mu e.
 */
static void local_apply_elastic2d(int n, const double *D, const double *jw, const double *dfi,
                                  double lam, double mu, const double *ue, double *out) {
  int nd = n * n;
  double *v = (double *)malloc(sizeof(double) * 4 * nd);
  for (int q = 0; q < nd; q++) {
    int i = q % n, j = q / n;
    double gr[2][2]; /* gr[m][b] = d u_m / d xi_b */
    for (int m = 0; m < 2; m++) {
      const double *um = ue + m * nd;
      double g0 = 0, g1 = 0;
      for (int l = 0; l < n; l++) {
        g0 += D[i * n + l] * um[l + n * j];
        g1 += D[j * n + l] * um[i + n * l];
      }
      gr[m][0] = g0; gr[m][1] = g1;
    }
    const double *Di = dfi + 4 * q; /* Di[a*2+b] = (DF^{-1})_{ab} */
    double du[2][2];
    for (int m = 0; m < 2; m++)
      for (int a = 0; a < 2; a++) du[m][a] = Di[0 * 2 + a] * gr[m][0] + Di[1 * 2 + a] * gr[m][1];
    double tr = du[0][0] + du[1][1];
    double sg[2][2];
    for (int m = 0; m < 2; m++)
      for (int a = 0; a < 2; a++) sg[m][a] = mu * (du[m][a] + du[a][m]) + (m == a ? lam * tr : 0.0);
    double s = jw[q];
    for (int m = 0; m < 2; m++)
      for (int b = 0; b < 2; b++)
        v[4 * q + 2 * m + b] = s * (Di[b * 2 + 0] * sg[m][0] + Di[b * 2 + 1] * sg[m][1]);
  }
  for (int q = 0; q < nd; q++) {
    int i = q % n, j = q / n;
    for (int m = 0; m < 2; m++) {
      double s = 0.0;
      for (int l = 0; l < n; l++) {
        s += D[l * n + i] * v[4 * (l + n * j) + 2 * m + 0];
        s += D[l * n + j] * v[4 * (i + n * l) + 2 * m + 1];
      }
      out[m * nd + q] = s;
    }
  }
  free(v);
}

/* y = beta*y + alpha*A u for elasticity (MltAddVectorH1,
 * Computation/ProdMatVectH1.cxx:1228-1276).  u, y: dim*N component-blocked. */
void or_stiffness_elastic(int dim, int r, long E, long N, const int *elem_nodes, const double *jw,
                          const double *dfinv, const double *lam, const double *mu,
                          const int *color, int ncolor, double alpha, const double *u,
                          double beta, double *y) {
  int n = r + 1, nd = dim == 2 ? n * n : n * n * n, dd = dim * dim;
  double x[OR_MAXN], w[OR_MAXN], D[OR_MAXN * OR_MAXN];
  or_gl(r, x, w);
  or_gl_deriv(r, x, D);
  if (beta == 0.0)
    for (long g = 0; g < dim * N; g++) y[g] = 0.0;
  else if (beta != 1.0)
    for (long g = 0; g < dim * N; g++) y[g] *= beta;
  for (int c = 0; c < ncolor; c++) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < E; e++) {
      if (color[e] != c) continue;
      double *ue = (double *)malloc(sizeof(double) * 2 * dim * nd);
      double *out = ue + dim * nd;
      for (int m = 0; m < dim; m++)
        for (int q = 0; q < nd; q++) ue[m * nd + q] = u[m * N + elem_nodes[e * nd + q]];
      if (dim == 2)
        local_apply_elastic2d(n, D, jw + e * nd, dfinv + e * nd * dd, lam[e], mu[e], ue, out);
      else
        local_apply_elastic(n, D, jw + e * nd, dfinv + e * nd * dd, lam[e], mu[e], ue, out);
      for (int m = 0; m < dim; m++)
        for (int q = 0; q < nd; q++) y[m * N + elem_nodes[e * nd + q]] += alpha * out[m * nd + q];
      free(ue);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Ricker pulse through the degree-24 GL interpolator                       */
/* ------------------------------------------------------------------------ */

/* primitive of the Ricker wavelet (Source/TimeSource.cxx:33-51) */
static double ricker_primitive(double f0, double t) {
  const double pi = 3.14159265358979323846;
  double T0 = 1.0 / f0;
  if ((t <= -0.9 * T0) || (t >= 2.9 * T0)) return 0.0;
  double tmp = f0 * t - 1.0;
  double pi2 = pi * pi;
  double g = exp(-(tmp * tmp) * pi2);
  return -tmp * g / f0;
}

/* closed-form Ricker (TimeSource.cxx:55-67), used only by tests */
double or_ricker(double f0, double t) {
  const double pi = 3.14159265358979323846;
  double T0 = 1.0 / f0;
  if ((t <= -0.9 * T0) || (t >= 2.9 * T0)) return 0.0;
  double tmp = f0 * t - 1.0;
  double a = (tmp * pi) * (tmp * pi);
  return (2.0 * a - 1.0) * exp(-a);
}

/*
 * Pulse values added at each leapfrog step.  Step n (loop time
 * t = t_begin + n*dt, ModifiedEquation.cxx:42) adds pulse(t+dt)*b
 * (ModifiedEquation.cxx:316) where pulse is the first derivative of the
 * degree-24 GL interpolant of the Ricker primitive over windows
 * [t0, t0 + 0.5/f0) that slide forward when t > t1 + eps
 * (TimeSourceHyperbolic::Init/GenerateValues/EvaluateDerivative,
 * Source/TimeSource.cxx:818-880; eps = (t_end - t_begin)*1e-15,
 * Instationary/VarInstationary.cxx:993), and zero for t >= 2.5/f0
 * (VarInstationary.cxx:416-420, 2394-2395).  t_end = t_begin + nsteps*dt.
 */
void or_pulse_table(double f0, double t_begin, double dt, int nsteps, double *pulse) {
  const int R = 24, NP = 25;
  double x[OR_MAXN1D], w[OR_MAXN1D], D[OR_MAXN1D * OR_MAXN1D], c[OR_MAXN1D];
  or_gl(R, x, w);
  or_gl_deriv(R, x, D);
  cte_phi(R, x, c);
  double t_end = t_begin + nsteps * dt;
  double eps = (t_end - t_begin) * 1e-15;
  double tlimit = 2.5 / f0;
  double t0 = t_begin, t1 = t0 + 0.5 / f0;
  double V0[OR_MAXN1D], V1[OR_MAXN1D];
#define OR_GEN_VALUES()                                                      \
  do {                                                                       \
    for (int i = 0; i < NP; i++) {                                           \
      double tt = (1.0 - x[i]) * t0 + x[i] * t1;                             \
      V0[i] = ricker_primitive(f0, tt);                                      \
    }                                                                        \
    for (int j = 0; j < NP; j++) {                                           \
      double s = 0.0;                                                        \
      for (int k = 0; k < NP; k++) s += D[j * NP + k] * V0[k];               \
      V1[j] = s * (1.0 / (t1 - t0));                                         \
    }                                                                        \
  } while (0)
  OR_GEN_VALUES();
  for (int n = 0; n < nsteps; n++) {
    double t = t_begin + n * dt;
    double ts = t + dt;
    if (ts >= tlimit) { pulse[n] = 0.0; continue; }
    if ((ts > t1 + eps) && (t1 > t0)) {
      while (ts > t1 + eps) {
        double tnext = t1 + (t1 - t0);
        t0 = t1; t1 = tnext;
      }
      OR_GEN_VALUES();
    }
    double p = 0.0;
    for (int i = 0; i < NP; i++) {
      double lambda = (ts - t0) / (t1 - t0);
      double ph = 1.0;
      for (int k = 0; k < NP; k++)
        if (k != i) ph *= lambda - x[k];
      ph *= c[i];
      p += ph * V1[i];
    }
    pulse[n] = p;
  }
#undef OR_GEN_VALUES
}

/* ------------------------------------------------------------------------ */
/* Point location, Dirac source, receivers                                  */
/* ------------------------------------------------------------------------ */

/* Newton inversion of the multilinear map of element e; returns 1 when the
 * reference point lies in [-tol, 1+tol]^dim, with xi clamped to [0,1]
 * (FjInverseProblem::Solve, FiniteElement/PointsReference.cxx:277-360, then
 * ProjectPointOnBoundary for points found slightly outside,
 * Source/DefineSourceElliptic.cxx:2539-2543). */
static int locate_in_elem(int dim, int nx, const double *vert, long e, const double *p,
                          double *xi) {
  long cv[8];
  elem_corners(dim, nx, e, cv);
  double lo[3], hi[3];
  for (int a = 0; a < dim; a++) { lo[a] = 1e300; hi[a] = -1e300; }
  for (int c = 0; c < (1 << dim); c++)
    for (int a = 0; a < dim; a++) {
      double v = vert[cv[c] * dim + a];
      if (v < lo[a]) lo[a] = v;
      if (v > hi[a]) hi[a] = v;
    }
  double h = hi[0] - lo[0];
  for (int a = 0; a < dim; a++) {
    double tol = 1e-10 * h;
    if (p[a] < lo[a] - tol || p[a] > hi[a] + tol) return 0;
  }
  for (int a = 0; a < dim; a++) xi[a] = 0.5;
  for (int it = 0; it < 50; it++) {
    double N8[8], dN[24], F[3], DF[9], Di[9];
    multilinear(dim, xi, N8, dN);
    for (int a = 0; a < dim; a++) {
      double s = 0.0;
      for (int c = 0; c < (1 << dim); c++) s += N8[c] * vert[cv[c] * dim + a];
      F[a] = s - p[a];
      for (int b = 0; b < dim; b++) {
        double d = 0.0;
        for (int c = 0; c < (1 << dim); c++) d += dN[c * dim + b] * vert[cv[c] * dim + a];
        DF[a * dim + b] = d;
      }
    }
    det_inv(dim, DF, Di);
    double nrm = 0.0;
    for (int a = 0; a < dim; a++) {
      double d = 0.0;
      for (int b = 0; b < dim; b++) d += Di[a * dim + b] * F[b];
      xi[a] -= d;
      nrm += d * d;
    }
    if (nrm < 1e-32) break;
  }
  const double tol = 1e-10;
  for (int a = 0; a < dim; a++)
    if (xi[a] < -tol || xi[a] > 1.0 + tol) return 0;
  for (int a = 0; a < dim; a++) {
    if (xi[a] < 0.0) xi[a] = 0.0;
    if (xi[a] > 1.0) xi[a] = 1.0;
  }
  return 1;
}

/* tensor Lagrange values phi_q(xi) of the Q_r element */
static void tensor_phi(int dim, int r, const double *x, const double *xi, double *phi) {
  int n = r + 1;
  double p1[3][OR_MAXN];
  for (int a = 0; a < dim; a++)
    for (int i = 0; i < n; i++) p1[a][i] = or_phi(r, x, i, xi[a]);
  int nd = (dim == 2) ? n * n : n * n * n;
  for (int q = 0; q < nd; q++) {
    int i = q % n, j = (q / n) % n, k = q / (n * n);
    double v = p1[0][i] * p1[1][j];
    if (dim == 3) v *= p1[2][k];
    phi[q] = v;
  }
}

/*
 * Dirac source vector b (full length N*ncomp, component-blocked) for a point
 * p with polarisation pol[ncomp]: b_i = pol_m phi_i(F_e^{-1}(p)) summed over
 * the elements containing p and divided by their number
 * (AddDiracSourceGen, Source/DefineSourceElliptic.cxx:2508-2725), Dirichlet
 * entries removed and entries below eps*||b||_2 dropped
 * (Instationary/VarInstationary.cxx:637-658).  Returns the element count.
 */
int or_dirac(int dim, int nx, int r, const double *vert, const int *elem_nodes, long N,
             const unsigned char *dirichlet, const double *p, int ncomp, const double *pol,
             double *b) {
  int n = r + 1;
  int nd = (dim == 2) ? n * n : n * n * n;
  long E = (dim == 2) ? (long)nx * nx : (long)nx * nx * nx;
  double x[OR_MAXN1D], w[OR_MAXN1D];
  or_gl(r, x, w);
  for (long g = 0; g < N * ncomp; g++) b[g] = 0.0;
  int cnt = 0;
  double phi[OR_MAXN * OR_MAXN * OR_MAXN];
  for (long e = 0; e < E; e++) {
    double xi[3];
    if (!locate_in_elem(dim, nx, vert, e, p, xi)) continue;
    tensor_phi(dim, r, x, xi, phi);
    for (int m = 0; m < ncomp; m++)
      for (int q = 0; q < nd; q++) b[m * N + elem_nodes[e * nd + q]] += pol[m] * phi[q];
    cnt++;
  }
  if (cnt > 1)
    for (long g = 0; g < N * ncomp; g++) b[g] *= 1.0 / cnt;
  double nrm = 0.0;
  for (long g = 0; g < N * ncomp; g++) {
    if (dirichlet[g % N]) b[g] = 0.0;
    nrm += b[g] * b[g];
  }
  double thr = 2.220446049250313e-16 * sqrt(nrm);
  for (long g = 0; g < N * ncomp; g++)
    if (fabs(b[g]) <= thr) b[g] = 0.0;
  return cnt;
}

/*
 * Receiver interpolation: first element containing p, weights phi_q(xi)
 * (Output/GridInterpolation.cxx point location + interpolation used by
 * SismoPoint / SismoCircle, :836-850, :1966-1983).  Returns element or -1.
 */
long or_receiver(int dim, int nx, int r, const double *vert, const double *p, double *wq) {
  long E = (dim == 2) ? (long)nx * nx : (long)nx * nx * nx;
  double x[OR_MAXN1D], w[OR_MAXN1D];
  or_gl(r, x, w);
  for (long e = 0; e < E; e++) {
    double xi[3];
    if (!locate_in_elem(dim, nx, vert, e, p, xi)) continue;
    tensor_phi(dim, r, x, xi, wq);
    return e;
  }
  return -1;
}

/* derivative of the Lagrange basis function i of the points x at t */
static double or_dphi(int r, const double *x, int i, double t) {
  double c[OR_MAXN1D];
  cte_phi(r, x, c);
  double s = 0.0;
  for (int l = 0; l <= r; l++) {
    if (l == i) continue;
    double p = 1.0;
    for (int k = 0; k <= r; k++)
      if (k != i && k != l) p *= t - x[k];
    s += p;
  }
  return s * c[i];
}

/*
 * Receiver value and gradient weights at p (the elastic SismoPoint output
 * interpolates u and grad u at the point, Output/GridInterpolation.cxx
 * :836-850 with compute_grad): wq[q] = phi_q(xi), wg[a*nd + q] =
 * d phi_q / d x_a = sum_b DF^{-1}[b][a] d phi_q / d xi_b with DF the Jacobian
 * of the element's multilinear map at xi.  Returns the element or -1.
 */
long or_receiver_grad(int dim, int nx, int r, const double *vert, const double *p, double *wq,
                      double *wg) {
  long E = (dim == 2) ? (long)nx * nx : (long)nx * nx * nx;
  int n = r + 1, nd = (dim == 2) ? n * n : n * n * n;
  double x[OR_MAXN1D], w[OR_MAXN1D];
  or_gl(r, x, w);
  for (long e = 0; e < E; e++) {
    double xi[3];
    if (!locate_in_elem(dim, nx, vert, e, p, xi)) continue;
    tensor_phi(dim, r, x, xi, wq);
    long cv[8];
    elem_corners(dim, nx, e, cv);
    double N8[8], dN[24], DF[9] = {0}, Di[9];
    multilinear(dim, xi, N8, dN);
    for (int a = 0; a < dim; a++)
      for (int b = 0; b < dim; b++) {
        double d = 0.0;
        for (int c = 0; c < (1 << dim); c++) d += dN[c * dim + b] * vert[cv[c] * dim + a];
        DF[a * dim + b] = d;
      }
    det_inv(dim, DF, Di);
    double v1[3][OR_MAXN1D], d1[3][OR_MAXN1D];
    for (int a = 0; a < dim; a++)
      for (int i = 0; i < n; i++) {
        v1[a][i] = or_phi(r, x, i, xi[a]);
        d1[a][i] = or_dphi(r, x, i, xi[a]);
      }
    for (int q = 0; q < nd; q++) {
      int idx[3] = {q % n, (q / n) % n, q / (n * n)};
      double gref[3];
      for (int b = 0; b < dim; b++) {
        double v = 1.0;
        for (int a = 0; a < dim; a++) v *= (a == b) ? d1[a][idx[a]] : v1[a][idx[a]];
        gref[b] = v;
      }
      for (int a = 0; a < dim; a++) {
        double g = 0.0;
        for (int b = 0; b < dim; b++) g += Di[b * dim + a] * gref[b];
        wg[a * nd + q] = g;
      }
    }
    return e;
  }
  return -1;
}

/* SismoCircle points (GridInterpolation.cxx:1966-1983 bounds, :2721-2740) */
void or_circle_points(double cx, double cy, double radius, int npts, double *pts) {
  const double pi = 3.14159265358979323846;
  double xmin = cx - radius, xmax = cx + radius, ymin = cy - radius, ymax = cy + radius;
  double step = (2.0 * pi) / npts;
  double c0 = (xmin + xmax) / 2.0, c1 = (ymin + ymax) / 2.0, rad = (xmax - xmin) / 2.0;
  for (int i = 0; i < npts; i++) {
    pts[2 * i] = c0 + rad * cos(step * i);
    pts[2 * i + 1] = c1 + rad * sin(step * i);
  }
}

/* ------------------------------------------------------------------------ */
/* Leapfrog                                                                 */
/* ------------------------------------------------------------------------ */

/*
 * Second-order leapfrog on D u'' + S u' + A u = f(t) b (S diagonal, NULL =
 * undamped) with homogeneous Dirichlet
 * nodes, restating RunSecondOrderScheme (Instationary/ModifiedEquation.cxx:
 * 16-53), SetInitialConditionS (:99-124, order 2) and Advance (:295-345,
 * order 2, S_h = 0):
 *   U0 = u0, U1 = U0 + dt du0 + dt^2/2 D^{-1}(-A U0 - S du0);  Dirichlet on both
 *   each step n (t = t_begin + n dt):  record R U0;
 *     Y = -A U1 + pulse[n] b;  Y = dt^2 Y - (D - dt/2 S) U0;  Y += 2 D U1;
 *     U0 = U1; U1 = Y; zero Dirichlet; U1 *= (D + dt/2 S)^{-1}
 *   after the loop record R U0  (GiveFinalIterate).
 * ncomp = 1 (acoustic, G) or dim (elastic, jw/dfinv/lam/mu).
 * Inputs u0/du0 may be NULL (zero).  Outputs: uN = U^nsteps, uN1 = U^{nsteps+1}
 * (the scheme's U1 at exit), traces[(nsteps+1)*nrec] = sum_q wr_q U0[node_q]
 * on component rcomp[rec].
 */
void or_leapfrog(int dim, int r, int ncomp, long E, long N, const int *elem_nodes,
                 const double *G, const double *jw, const double *dfinv, const double *lam,
                 const double *mu, const int *color, int ncolor, const double *Dm, const double *Sd,
                 const unsigned char *dirichlet, const double *b, const double *pulse,
                 double dt, int nsteps, const double *u0, const double *du0, int nrec,
                 const long *rec_elem, const double *rec_w, const int *rec_comp,
                 double *uN, double *uN1, double *traces) {
  int n = r + 1, nd = (dim == 2) ? n * n : n * n * n;
  long NT = N * ncomp;
  double *U0 = (double *)calloc((size_t)NT, sizeof(double));
  double *U1 = (double *)calloc((size_t)NT, sizeof(double));
  double *Y = (double *)calloc((size_t)NT, sizeof(double));
  double *invD = (double *)malloc(sizeof(double) * NT);
  double *invDp = (double *)malloc(sizeof(double) * NT);
  double *Dmin = (double *)malloc(sizeof(double) * NT);
  for (long g = 0; g < NT; g++) invD[g] = 1.0 / Dm[g % N];
  /* damping: D - dt/2 S and (D + dt/2 S)^{-1} (TimeAcoustic.cxx:977-1004) */
  for (long g = 0; g < NT; g++) {
    double sg = Sd ? Sd[g % N] : 0.0;
    Dmin[g] = Dm[g % N] - 0.5 * dt * sg;
    invDp[g] = 1.0 / (Dm[g % N] + 0.5 * dt * sg);
  }
  double dt2 = dt * dt;
  long nb = 0;
  for (long g = 0; g < NT; g++)
    if (b && b[g] != 0.0) nb++;
  long *bi = (long *)malloc(sizeof(long) * (nb + 1));
  nb = 0;
  for (long g = 0; g < NT; g++)
    if (b && b[g] != 0.0) bi[nb++] = g;
#define OR_APPLY(alpha, uu, yy)                                                            \
  do {                                                                                     \
    if (ncomp == 1)                                                                        \
      or_stiffness(dim, r, E, N, elem_nodes, G, color, ncolor, (alpha), (uu), 0.0, (yy));  \
    else                                                                                   \
      or_stiffness_elastic(dim, r, E, N, elem_nodes, jw, dfinv, lam, mu, color, ncolor, (alpha), \
                           (uu), 0.0, (yy));                                               \
  } while (0)
#define OR_DIRICHLET(vv)                                  \
  do {                                                    \
    for (long g = 0; g < NT; g++)                         \
      if (dirichlet[g % N]) (vv)[g] = 0.0;                \
  } while (0)
#define OR_RECORD(row)                                                          \
  do {                                                                          \
    for (int s = 0; s < nrec; s++) {                                            \
      double acc = 0.0;                                                         \
      long e = rec_elem[s];                                                     \
      int m = rec_comp ? rec_comp[s] : 0;                                       \
      for (int q = 0; q < nd; q++)                                              \
        acc += rec_w[(long)s * nd + q] * U0[m * N + elem_nodes[e * nd + q]];    \
      traces[(long)(row)*nrec + s] = acc;                                       \
    }                                                                           \
  } while (0)
  if (u0) memcpy(U0, u0, sizeof(double) * NT);
  memcpy(U1, U0, sizeof(double) * NT);
  OR_DIRICHLET(U0);
  /* U1 = D^{-1}(K U0 - S du0) with K = -A */
  OR_APPLY(-1.0, U0, Y);
  if (Sd && du0)
    for (long g = 0; g < NT; g++) Y[g] -= Sd[g % N] * du0[g];
  for (long g = 0; g < NT; g++) U1[g] = Y[g] * invD[g];
  for (long g = 0; g < NT; g++) {
    double v = U1[g] * (0.5 * dt);
    v += du0 ? du0[g] : 0.0;
    v *= dt;
    U1[g] = v + U0[g];
  }
  OR_DIRICHLET(U1);
  for (int it = 0; it < nsteps; it++) {
    OR_RECORD(it);
    OR_APPLY(-1.0, U1, Y);
    if (pulse && pulse[it] != 0.0)
      for (long s = 0; s < nb; s++) Y[bi[s]] += pulse[it] * b[bi[s]];
    for (long g = 0; g < NT; g++) Y[g] = dt2 * Y[g] + (-1.0) * (Dmin[g] * U0[g]);
    for (long g = 0; g < NT; g++) Y[g] = 1.0 * Y[g] + 2.0 * (Dm[g % N] * U1[g]);
    double *tmp = U0; U0 = U1; U1 = Y; Y = tmp;
    OR_DIRICHLET(U1);
    for (long g = 0; g < NT; g++) U1[g] *= invDp[g];
  }
  OR_RECORD(nsteps);
  if (uN) memcpy(uN, U0, sizeof(double) * NT);
  if (uN1) memcpy(uN1, U1, sizeof(double) * NT);
#undef OR_APPLY
#undef OR_DIRICHLET
#undef OR_RECORD
  free(U0); free(U1); free(Y); free(invD); free(invDp); free(Dmin); free(bi);
}

/*
 * Fourth-order modified-equation leapfrog (ModifiedEquationIterator, order
 * 4), step for step as Montjoie runs it with Kh = -A:
 * SetInitialConditionS (Instationary/ModifiedEquation.cxx:124-170): from
 * U1 = u0 (Dirichlet rows 0) and du0, d2 = D^{-1}(Kh U1 - S du0),
 * d3 = D^{-1}(Kh du0 - S d2), d4 = D^{-1}(Kh d2 - S d3) (each with Dirichlet
 * rows 0), then U2 = U1 + dt du0 + dt^2/2 d2 + dt^3/6 d3 + dt^4/24 d4 and U0
 * the same with -dt, in Montjoie's Horner order.  Advance (:347-407) at step
 * it (t = it*dt):
 *   P  = dt^2 (Kh U2 + f(t+dt) b)
 *   V  = dt/2 (U0 - 4 U1 + 3 U2)
 *   W  = D^{-1}(P - S V), Dirichlet 0
 *   X  = D^{-1}(Kh V - S W + dt^2 f'(t+dt) b), Dirichlet 0
 *   Z  = Kh W - S X + dt^2 f''(t+dt) b
 *   Q  = P + dt^2/12 Z + dt^2/6 S X + 2 D U2 - (D - dt/2 S) U1
 *   U0 <- U1, U1 <- U2, U2 <- (D + dt/2 S)^{-1} Q, Dirichlet 0.
 * The iterate of step it is U1 (GetIterate): traces row it = R U1 before the
 * step, row nsteps after the last.  pulse, d1, d2: f, f', f'' at t_it + dt
 * (or_source_table with nb_deriv 1, 2, 3), each may be NULL (zero).
 * uN, uN1 receive U1 and U2 at the end.
 */
void or_leapfrog4(int dim, int r, int ncomp, long E, long N, const int *elem_nodes,
                  const double *G, const double *jw, const double *dfinv, const double *lam,
                  const double *mu, const int *color, int ncolor, const double *Dm,
                  const double *Sd, const unsigned char *dirichlet, const double *b,
                  const double *pulse, const double *d1, const double *d2, double dt,
                  int nsteps, const double *u0, const double *du0, int nrec,
                  const long *rec_elem, const double *rec_w, const int *rec_comp, double *uN,
                  double *uN1, double *traces) {
  int n = r + 1, nd = (dim == 2) ? n * n : n * n * n;
  long NT = N * ncomp;
  size_t bytes = sizeof(double) * (size_t)NT;
  double *U0 = (double *)calloc((size_t)NT, sizeof(double));
  double *U1 = (double *)calloc((size_t)NT, sizeof(double));
  double *U2 = (double *)calloc((size_t)NT, sizeof(double));
  double *V = (double *)calloc((size_t)NT, sizeof(double));
  double *P = (double *)calloc((size_t)NT, sizeof(double));
  double *W = (double *)calloc((size_t)NT, sizeof(double));
  double *X = (double *)calloc((size_t)NT, sizeof(double));
  double *Z = (double *)calloc((size_t)NT, sizeof(double));
  double dt2 = dt * dt;
#define OR4_APPLY(uu, yy)                                                                   \
  do {                                                                                      \
    if (ncomp == 1)                                                                         \
      or_stiffness(dim, r, E, N, elem_nodes, G, color, ncolor, -1.0, (uu), 0.0, (yy));      \
    else                                                                                    \
      or_stiffness_elastic(dim, r, E, N, elem_nodes, jw, dfinv, lam, mu, color, ncolor, -1.0, \
                           (uu), 0.0, (yy));                                                \
  } while (0)
#define OR4_S(g) (Sd ? Sd[(g) % N] : 0.0)
#define OR4_DIR(vv)                                         \
  do {                                                      \
    for (long g = 0; g < NT; g++)                           \
      if (dirichlet[g % N]) (vv)[g] = 0.0;                  \
  } while (0)
#define OR4_SRC(tab, it, scale, yy)                          \
  do {                                                       \
    if ((tab) && (tab)[it] != 0.0 && b)                      \
      for (long g = 0; g < NT; g++)                          \
        if (b[g] != 0.0) (yy)[g] += (scale) * ((tab)[it] * b[g]); \
  } while (0)
#define OR4_RECORD(row)                                                         \
  do {                                                                          \
    for (int s = 0; s < nrec; s++) {                                            \
      double acc = 0.0;                                                         \
      long e = rec_elem[s];                                                     \
      int m = rec_comp ? rec_comp[s] : 0;                                       \
      for (int q = 0; q < nd; q++)                                              \
        acc += rec_w[(long)s * nd + q] * U1[m * N + elem_nodes[e * nd + q]];    \
      traces[(long)(row)*nrec + s] = acc;                                       \
    }                                                                           \
  } while (0)
  /* initial condition */
  if (u0) memcpy(U1, u0, bytes);
  OR4_DIR(U1);
  if (du0) memcpy(V, du0, bytes);
  double *d2v = W, *d3v = X, *d4v = Z;
  OR4_APPLY(U1, d2v);
  for (long g = 0; g < NT; g++) d2v[g] = (d2v[g] - OR4_S(g) * V[g]) / Dm[g % N];
  OR4_DIR(d2v);
  OR4_APPLY(V, d3v);
  for (long g = 0; g < NT; g++) d3v[g] = (d3v[g] - OR4_S(g) * d2v[g]) / Dm[g % N];
  OR4_DIR(d3v);
  OR4_APPLY(d2v, d4v);
  for (long g = 0; g < NT; g++) d4v[g] = (d4v[g] - OR4_S(g) * d3v[g]) / Dm[g % N];
  OR4_DIR(d4v);
  for (long g = 0; g < NT; g++) {
    double a4 = d4v[g] / 24.0, a3 = d3v[g] / 6.0, a2 = d2v[g] / 2.0;
    double up = a4 * dt; up += a3; up *= dt; up += a2; up *= dt; up += V[g]; up *= dt;
    U2[g] = up + U1[g];
    double um = a4 * dt; um -= a3; um *= dt; um += a2; um *= dt; um -= V[g]; um *= dt;
    U0[g] = um + U1[g];
  }
  OR4_DIR(U1);
  OR4_DIR(U2);
  for (int it = 0; it < nsteps; it++) {
    OR4_RECORD(it);
    OR4_APPLY(U2, P);
    OR4_SRC(pulse, it, 1.0, P);
    for (long g = 0; g < NT; g++) P[g] *= dt2;
    for (long g = 0; g < NT; g++) V[g] = 0.5 * dt * (U0[g] - 4.0 * U1[g] + 3.0 * U2[g]);
    for (long g = 0; g < NT; g++) W[g] = (P[g] - OR4_S(g) * V[g]) / Dm[g % N];
    OR4_DIR(W);
    OR4_APPLY(V, X);
    for (long g = 0; g < NT; g++) X[g] -= OR4_S(g) * W[g];
    OR4_SRC(d1, it, dt2, X);
    for (long g = 0; g < NT; g++) X[g] /= Dm[g % N];
    OR4_DIR(X);
    OR4_APPLY(W, Z);
    for (long g = 0; g < NT; g++) Z[g] -= OR4_S(g) * X[g];
    OR4_SRC(d2, it, dt2, Z);
    for (long g = 0; g < NT; g++) {
      double q = P[g] + dt2 / 12.0 * Z[g];
      q += dt2 / 6.0 * (OR4_S(g) * X[g]);
      q += 2.0 * (Dm[g % N] * U2[g]);
      q -= (Dm[g % N] - 0.5 * dt * OR4_S(g)) * U1[g];
      U0[g] = q / (Dm[g % N] + 0.5 * dt * OR4_S(g)); /* the new level, in U0's storage */
    }
    double *tmp = U0; U0 = U1; U1 = U2; U2 = tmp;
    OR4_DIR(U2);
  }
  OR4_RECORD(nsteps);
  if (uN) memcpy(uN, U1, bytes);
  if (uN1) memcpy(uN1, U2, bytes);
#undef OR4_APPLY
#undef OR4_S
#undef OR4_DIR
#undef OR4_SRC
#undef OR4_RECORD
  free(U0); free(U1); free(U2); free(V); free(P); free(W); free(X); free(Z);
}

/* ------------------------------------------------------------------------ */
/* Largest eigenvalue of D^{-1} A (CFL)                                     */
/* ------------------------------------------------------------------------ */

/* largest eigenvalue of a symmetric tridiagonal (alpha, beta) by bisection */
static double tridiag_max_eig(int m, const double *al, const double *be) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; i++) {
    double rad = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < m - 1 ? fabs(be[i]) : 0.0);
    if (al[i] - rad < lo) lo = al[i] - rad;
    if (al[i] + rad > hi) hi = al[i] + rad;
  }
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    int cnt = 0; /* number of eigenvalues < mid (Sturm count) */
    double d = 1.0;
    for (int i = 0; i < m; i++) {
      d = al[i] - mid - (i > 0 ? be[i - 1] * be[i - 1] / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) cnt++;
    }
    if (cnt == m) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

/*
 * lambda_max(D^{-1} A) on the non-Dirichlet nodes by Lanczos on
 * D^{-1/2} A D^{-1/2} (FP64, full reorthogonalisation), the quantity behind
 * dt_max = 2 / sqrt(lambda_max) (EvaluateCFL_Generic,
 * Instationary/EvaluationCflWave.cxx:717-760, P(0) = 2 from
 * ModifiedEquation.cxx:217-225; the reference's ARPACK call is replaced).
 */
double or_lambda_max(int dim, int r, int ncomp, long E, long N, const int *elem_nodes,
                     const double *G, const double *jw, const double *dfinv, const double *lam,
                     const double *mu, const int *color, int ncolor, const double *Dm,
                     const unsigned char *dirichlet, int m, uint64_t seed) {
  long NT = N * ncomp;
  double *Q = (double *)calloc((size_t)NT * (m + 1), sizeof(double));
  double *wv = (double *)malloc(sizeof(double) * NT);
  double *t1 = (double *)malloc(sizeof(double) * NT);
  double *sq = (double *)malloc(sizeof(double) * NT);
  double *al = (double *)calloc(m, sizeof(double)), *be = (double *)calloc(m, sizeof(double));
  for (long g = 0; g < NT; g++) sq[g] = dirichlet[g % N] ? 0.0 : 1.0 / sqrt(Dm[g % N]);
  double nrm = 0.0;
  for (long g = 0; g < NT; g++) {
    double v = dirichlet[g % N] ? 0.0 : or_uniform(seed, g) - 0.5;
    Q[g] = v;
    nrm += v * v;
  }
  nrm = sqrt(nrm);
  for (long g = 0; g < NT; g++) Q[g] /= nrm;
  int k;
  for (k = 0; k < m; k++) {
    double *q = Q + (long)k * NT;
    for (long g = 0; g < NT; g++) t1[g] = q[g] * sq[g];
    if (ncomp == 1)
      or_stiffness(dim, r, E, N, elem_nodes, G, color, ncolor, 1.0, t1, 0.0, wv);
    else
      or_stiffness_elastic(dim, r, E, N, elem_nodes, jw, dfinv, lam, mu, color, ncolor, 1.0, t1, 0.0, wv);
    for (long g = 0; g < NT; g++) wv[g] *= sq[g];
    double a = 0.0;
    for (long g = 0; g < NT; g++) a += wv[g] * q[g];
    al[k] = a;
    for (int pass = 0; pass < 2; pass++)
      for (int j = 0; j <= k; j++) {
        double *qj = Q + (long)j * NT, d = 0.0;
        for (long g = 0; g < NT; g++) d += wv[g] * qj[g];
        for (long g = 0; g < NT; g++) wv[g] -= d * qj[g];
      }
    double bn = 0.0;
    for (long g = 0; g < NT; g++) bn += wv[g] * wv[g];
    bn = sqrt(bn);
    if (k < m - 1) be[k] = bn;
    if (bn < 1e-300) { k++; break; }
    double *qn = Q + (long)(k + 1) * NT;
    for (long g = 0; g < NT; g++) qn[g] = wv[g] / bn;
  }
  double lmax = tridiag_max_eig(k, al, be);
  free(Q); free(wv); free(t1); free(sq); free(al); free(be);
  return lmax;
}

/* ------------------------------------------------------------------------ */
/* Adjoint gradient (FP64 reference)                                        */
/* ------------------------------------------------------------------------ */

/*
 * Gradient of a receiver misfit with respect to per-element mu for the
 * acoustic leapfrog with zero initial state.  No counterpart in the reference
 * (it has no time-domain adjoint); derived from the leapfrog recurrence of
 * ModifiedEquation.cxx:295-345 written as
 *   U^{k+1} = 2U^k - U^{k-1} + S(-A U^k + f_{k-1} b),  S = dt^2 D^{-1} (0 on
 *   Dirichlet rows), U^0 = U^1 = 0,  traces d^k = R U^k, k = 0..N.
 * With h^k = R^T dJ/dd^k the scaled multipliers nu satisfy
 *   nu_N = 0, nu_{N-1} = S h^N, nu_{j-1} = 2nu_j - nu_{j+1} + S(-A nu_j + h^j),
 * and dJ/dmu_e = -sum_{j=1}^{N-1} nu_j^T (A_e/mu_e) U^j.  All forward states
 * are stored (small problems only).  Pinned by central finite differences of
 * the misfit in tests/test_oracle.py.
 */
void or_adjoint_acoustic(int dim, int r, long E, long N, const int *elem_nodes, const double *G,
                         const double *mu, const int *color, int ncolor, const double *Dm,
                         const unsigned char *dirichlet, const double *b, const double *pulse,
                         double dt, int nsteps, int nrec, const long *rec_elem,
                         const double *rec_w, const double *dJ, double *grad) {
  int n = r + 1, nd = (dim == 2) ? n * n : n * n * n;
  int nc = (dim == 2) ? 3 : 6;
  double *U = (double *)calloc((size_t)N * (nsteps + 1), sizeof(double));
  double *S = (double *)malloc(sizeof(double) * N);
  double *t1 = (double *)malloc(sizeof(double) * N);
  double *nu0 = (double *)calloc(N, sizeof(double)), *nu1 = (double *)calloc(N, sizeof(double));
  double *h = (double *)malloc(sizeof(double) * N);
  for (long g = 0; g < N; g++) S[g] = dirichlet[g] ? 0.0 : dt * dt / Dm[g];
  for (int k = 1; k < nsteps; k++) {
    double *Uk = U + (long)k * N, *Ukm = U + (long)(k - 1) * N, *Ukp = U + (long)(k + 1) * N;
    or_stiffness(dim, r, E, N, elem_nodes, G, color, ncolor, 1.0, Uk, 0.0, t1);
    for (long g = 0; g < N; g++) {
      double f = -t1[g] + (b ? pulse[k - 1] * b[g] : 0.0);
      Ukp[g] = 2.0 * Uk[g] - Ukm[g] + S[g] * f;
    }
  }
#define OR_HVEC(row)                                                          \
  do {                                                                        \
    for (long g = 0; g < N; g++) h[g] = 0.0;                                  \
    for (int s = 0; s < nrec; s++) {                                          \
      long e = rec_elem[s];                                                   \
      for (int q = 0; q < nd; q++)                                            \
        h[elem_nodes[e * nd + q]] += rec_w[(long)s * nd + q] * dJ[(long)(row)*nrec + s]; \
    }                                                                         \
  } while (0)
  /* nu1 = nu_{j+1}, nu0 = nu_j; start j = N-1 */
  for (long g = 0; g < N; g++) nu1[g] = 0.0;
  OR_HVEC(nsteps);
  for (long g = 0; g < N; g++) nu0[g] = S[g] * h[g];
  for (long e = 0; e < E; e++) grad[e] = 0.0;
  double ue[OR_MAXN * OR_MAXN * OR_MAXN], ve[OR_MAXN * OR_MAXN * OR_MAXN],
      oe[OR_MAXN * OR_MAXN * OR_MAXN];
  double x[OR_MAXN1D], w[OR_MAXN1D], D[OR_MAXN1D * OR_MAXN1D];
  or_gl(r, x, w);
  or_gl_deriv(r, x, D);
  for (int j = nsteps - 1; j >= 1; j--) {
    const double *Uj = U + (long)j * N;
    /* gradient term -nu_j^T A_e^0 U^j */
    for (long e = 0; e < E; e++) {
      for (int q = 0; q < nd; q++) {
        ue[q] = Uj[elem_nodes[e * nd + q]];
        ve[q] = nu0[elem_nodes[e * nd + q]];
      }
      local_apply(dim, n, D, G + e * nd * nc, ve, oe);
      double s = 0.0;
      for (int q = 0; q < nd; q++) s += oe[q] * ue[q];
      grad[e] -= s / mu[e];
    }
    if (j == 1) break;
    or_stiffness(dim, r, E, N, elem_nodes, G, color, ncolor, 1.0, nu0, 0.0, t1);
    OR_HVEC(j);
    for (long g = 0; g < N; g++) {
      double nm = 2.0 * nu0[g] - nu1[g] + S[g] * (-t1[g] + h[g]);
      nu1[g] = nu0[g];
      nu0[g] = nm;
    }
  }
#undef OR_HVEC
  free(U); free(S); free(t1); free(nu0); free(nu1); free(h);
}

/* ------------------------------------------------------------------------ */
/* Time table of the source, any derivative and offset                      */
/* ------------------------------------------------------------------------ */

/*
 * out[n] = f^(nb_deriv)(t_begin + n*dt + offset) for n < nsteps: the GL-24
 * interpolant of the Ricker primitive and its derivatives by repeated
 * application of the GL differentiation matrix (TimeSourceHyperbolic::
 * GenerateValues / EvaluateDerivative, Source/TimeSource.cxx:826-880), zero
 * from t = 2.5/f0 on (AddScalarSourceAtTime, VarInstationary.cxx:2394).  The
 * second-order leapfrog adds nb_deriv = 1 at offset dt (ModifiedEquation.cxx:
 * 316, or_pulse_table); the first-order system leapfrog adds nb_deriv = 0 at
 * offset dt/2 (ModifiedEquation.cxx:784-785).  nb_deriv <= 3.
 */
void or_time_table(double f0, double t_begin, double dt, int nsteps, int nb_deriv,
                   double offset, double *out) {
  const int R = 24, NP = 25;
  double x[OR_MAXN1D], w[OR_MAXN1D], D[OR_MAXN1D * OR_MAXN1D], c[OR_MAXN1D];
  or_gl(R, x, w);
  or_gl_deriv(R, x, D);
  cte_phi(R, x, c);
  double eps = (nsteps * dt) * 1e-15;
  double tlimit = 2.5 / f0;
  double t0 = t_begin, t1 = t0 + 0.5 / f0;
  double V[4][OR_MAXN1D];
  if (nb_deriv < 0 || nb_deriv > 3) return;
#define OR_GEN_DERIVS()                                                      \
  do {                                                                       \
    for (int i = 0; i < NP; i++)                                             \
      V[0][i] = ricker_primitive(f0, (1.0 - x[i]) * t0 + x[i] * t1);         \
    for (int d = 1; d <= nb_deriv; d++)                                      \
      for (int j = 0; j < NP; j++) {                                         \
        double s = 0.0;                                                      \
        for (int k = 0; k < NP; k++) s += D[j * NP + k] * V[d - 1][k];       \
        V[d][j] = s * (1.0 / (t1 - t0));                                     \
      }                                                                      \
  } while (0)
  OR_GEN_DERIVS();
  for (int n = 0; n < nsteps; n++) {
    double ts = (t_begin + n * dt) + offset;
    if (ts >= tlimit) { out[n] = 0.0; continue; }
    if ((ts > t1 + eps) && (t1 > t0)) {
      while (ts > t1 + eps) {
        double tnext = t1 + (t1 - t0);
        t0 = t1; t1 = tnext;
      }
      OR_GEN_DERIVS();
    }
    double p = 0.0, lambda = (ts - t0) / (t1 - t0);
    for (int i = 0; i < NP; i++) {
      double ph = 1.0;
      for (int k = 0; k < NP; k++)
        if (k != i) ph *= lambda - x[k];
      p += (ph * c[i]) * V[nb_deriv][i];
    }
    out[n] = p;
  }
#undef OR_GEN_DERIVS
}

/* ------------------------------------------------------------------------ */
/* Every TemporalSource kind of the explicit path                           */
/* ------------------------------------------------------------------------ */

/* cut-off time of each kind (tlimit_source, GetNewTimeSource,
 * Instationary/VarInstationary.cxx:409-481): 1 RICKER, 2 DERIVATIVE_RICKER,
 * 3 MODIFIED_RICKER (par = centre time), 4 GAUSSIAN (par = plateau width),
 * 5 HARMONIC (par = final time), 6 SINUS_GAUSSIAN (par = alpha) */
double or_source_limit(int kind, double f0, double par) {
  switch (kind) {
    case 1: case 2: return 2.5 / f0;
    case 3: return (2.5 + par) / (sqrt(2.0) * f0);
    case 4: return 2.5 / f0 + par;
    case 5: return par;
    case 6: return 2.0 * sqrt(log(1e6) / par);
  }
  return 0.0;
}

/* the function each kind interpolates (VirtualTimeSource::Evaluate of
 * Source/TimeSource.cxx: TimeRickerSource :33-51, DerivativeTimeRickerSource
 * :121-124 -> :55-67, TimeModifiedRickerSource :152-160, TimeGaussianSource
 * :198-213 with t_begin 0, TimeHarmonicSource :597-611,
 * TimeSinusGaussianSource :667-679 with toff half the cut-off) */
static double source_primitive(int kind, double f0, double par, double t) {
  const double pi = 3.14159265358979323846;
  if (kind == 1) return ricker_primitive(f0, t);
  if (kind == 2) return or_ricker(f0, t);
  if (kind == 3) {
    double u = f0 * (t - par);
    return -(t - par) * exp(-2.0 * (u * pi) * (u * pi));
  }
  if (kind == 4) {
    double T0 = 1.0 / f0, s;
    if (t < T0) s = t;
    else if (t >= T0 + par) s = t - par;
    else return 1.0;
    double u = pi * (f0 * s - 1.0);
    return exp(-u * u);
  }
  if (kind == 5) {
    if (t <= 0.0 || t >= par) return 0.0;
    double a = f0 * f0 * t * t, b = f0 * f0 * (t - par) * (t - par);
    return exp(-1.0 / (pi * a)) * sin(2.0 * pi * f0 * t) * exp(-1.0 / (pi * b));
  }
  if (kind == 6) {
    double toff = 0.5 * or_source_limit(6, f0, par), s = t - toff;
    return sin(2.0 * pi * f0 * t) * exp(-par * (s * s));
  }
  return 0.0;
}

/*
 * or_time_table for any kind: out[n] = f^(nb_deriv)(t_begin + n*dt +
 * offset), the GL-24 interpolant of the kind's function on windows of
 * 0.5/f0 (TimeSourceHyperbolic, Source/TimeSource.cxx:818-880), zero from
 * the kind's cut-off on (AddScalarSourceAtTime, VarInstationary.cxx:2394).
 * nb_deriv <= 3.
 */
void or_source_table(int kind, double f0, double par, double t_begin, double dt, int nsteps,
                     int nb_deriv, double offset, double *out) {
  const int R = 24, NP = 25;
  double x[OR_MAXN1D], w[OR_MAXN1D], D[OR_MAXN1D * OR_MAXN1D], c[OR_MAXN1D];
  or_gl(R, x, w);
  or_gl_deriv(R, x, D);
  cte_phi(R, x, c);
  double eps = (nsteps * dt) * 1e-15;
  double tlimit = or_source_limit(kind, f0, par);
  double t0 = t_begin, t1 = t0 + 0.5 / f0;
  double V[4][OR_MAXN1D];
  if (nb_deriv < 0 || nb_deriv > 3) return;
  int fresh = 1;
  for (int n = 0; n < nsteps; n++) {
    double ts = (t_begin + n * dt) + offset;
    if (ts >= tlimit) { out[n] = 0.0; continue; }
    if ((ts > t1 + eps) && (t1 > t0)) {
      while (ts > t1 + eps) {
        double tnext = t1 + (t1 - t0);
        t0 = t1; t1 = tnext;
      }
      fresh = 1;
    }
    if (fresh) {
      for (int i = 0; i < NP; i++)
        V[0][i] = source_primitive(kind, f0, par, (1.0 - x[i]) * t0 + x[i] * t1);
      for (int d = 1; d <= nb_deriv; d++)
        for (int j = 0; j < NP; j++) {
          double s = 0.0;
          for (int k = 0; k < NP; k++) s += D[j * NP + k] * V[d - 1][k];
          V[d][j] = s * (1.0 / (t1 - t0));
        }
      fresh = 0;
    }
    double p = 0.0, lambda = (ts - t0) / (t1 - t0);
    for (int i = 0; i < NP; i++) {
      double ph = 1.0;
      for (int k = 0; k < NP; k++)
        if (k != i) ph *= lambda - x[k];
      p += (ph * c[i]) * V[nb_deriv][i];
    }
    out[n] = p;
  }
}

/* ------------------------------------------------------------------------ */
/* First-order split system with PML layers (LEAP_FROG_PML)                 */
/* ------------------------------------------------------------------------ */

/* x = P^{-1} y for a dense k x k block (k <= 3) by Gaussian elimination
 * with partial pivoting */
static void or_block_solve(int k, const double *P, const double *y, double *x) {
  double A[3][4];
  for (int i = 0; i < k; i++) {
    for (int j = 0; j < k; j++) A[i][j] = P[i * k + j];
    A[i][k] = y[i];
  }
  for (int c = 0; c < k; c++) {
    int piv = c;
    for (int i = c + 1; i < k; i++)
      if (fabs(A[i][c]) > fabs(A[piv][c])) piv = i;
    if (piv != c)
      for (int j = 0; j <= k; j++) { double t = A[c][j]; A[c][j] = A[piv][j]; A[piv][j] = t; }
    for (int i = c + 1; i < k; i++) {
      double f = A[i][c] / A[c][c];
      for (int j = c; j <= k; j++) A[i][j] -= f * A[c][j];
    }
  }
  for (int i = k - 1; i >= 0; i--) {
    double s = A[i][k];
    for (int j = i + 1; j < k; j++) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
}

/*
 * Montjoie's first-order split acoustic system with PML layers, restated
 * for the structured (extended) mesh (dim, nx, r, vert):
 *   D dU/dt = R^S V - S U + F(t) b,     B dV/dt = R^V U - S^V V
 * with U = (u, u*, u**) -- u on every node, the split copies u* = u_x - u_y
 * (and u** = u_x - u_z in 3-D) on the nodes of PML elements -- and V (dim
 * values) at the quadrature points of every element (element matrices of
 * VarHelmholtz_Base::ComputeElementaryMatrixHelm, first-order branch,
 * Elliptic/Helmholtz/VarHelmholtz.cxx:296-378,620-660 with
 * HelmholtzElementaryMatrixClass::ComputeMatrixV / ComputeMatrixVdiag,
 * :3277-3440,3444-3720, coefficients :133-230; extracted into R^S, R^V by
 * VarInstationary_Base::ExtractStiffnessMatrix, Instationary/
 * VarInstationary.cxx:1423-1470; mass and damping blocks
 * Hyperbolic/Acoustic/TimeAcoustic.cxx:1079-1235):
 *   (R^S V)_i   = -sum_k w_k (J DF^{-T} grad phi_i)(xi_k) . v_k,
 *   (R^S V)*_i  = same with v_k replaced by (v_x, -v_y[, 0]),
 *   (R^S V)**_i = same with (v_x, 0, -v_z);
 *   (R^V U)_k   =  w_k J DF^{-T} grad u(xi_k)  (main u only);
 *   D = rho J w lumped (u: every element, u*, u**: PML elements only;
 *       Dirichlet rows of u = 1);
 *   S: u: sigma J w outside the PML; in a PML element, per node
 *       2-D [[s+, s-], [s-, s+]], s+ = rho (sx+sy)/2 Jw, s- = rho (sx-sy)/2 Jw;
 *       3-D rows (u, u*, u**): [S, d1 - d3, d2 + d3], [d1, S + d3, d1],
 *       [d2, d2, S - d3] with S = rho (sx+sy+sz)/3 Jw, d1 = rho (sx-sy)/3 Jw,
 *       d2 = rho (sx-sz)/3 Jw, d3 = rho (sy-sz)/3 Jw; Dirichlet rows 0;
 *   B = J w / mu, S^V_a = tau_a J w / mu (per component).
 * Scheme (ModifiedEquationSystemIterator, order 2, ModifiedEquation.cxx:
 * 511-560 with zero initial data, Advance :773-812, loop RunFirstOrderScheme
 * :441-481): V^{1/2} = 0; for n: record u^n;
 *   (D + dt/2 S) U^{n+1} = (D - dt/2 S) U^n + dt (R^S V^{n+1/2} + f_n b),
 *   u^{n+1} = 0 on Dirichlet nodes,
 *   (B + dt/2 S^V) V^{n+3/2} = (B - dt/2 S^V) V^{n+1/2} + dt R^V U^{n+1};
 * record u^nsteps.  f_n = or_time_table(nb_deriv 0, offset dt/2).
 * pml[e] != 0 marks PML elements, tau[(e*nd + q)*dim + a] the damping at
 * their quadrature points (GL nodes); sigma (per element, NULL = 0) the
 * physical damping outside the PML.  Outputs: uN[N] = u^nsteps,
 * traces[(nsteps+1)*nrec] of u as in or_leapfrog.
 */
void or_leapfrog_pml(int dim, int nx, int r, const double *vert, long N, const int *elem_nodes,
                     const double *rho, const double *mu, const double *sigma,
                     const unsigned char *pml, const double *tau,
                     const unsigned char *dirichlet, const double *b, const double *f, double dt,
                     int nsteps, int nrec, const long *rec_elem, const double *rec_w,
                     double *uN, double *traces) {
  int n = r + 1, nd = (dim == 2) ? n * n : n * n * n, d = dim;
  long E = (dim == 2) ? (long)nx * nx : (long)nx * nx * nx;
  double xg[OR_MAXN], wg[OR_MAXN], Dm[OR_MAXN * OR_MAXN];
  or_gl(r, xg, wg);
  or_gl_deriv(r, xg, Dm);
  /* per element node: Gk = w J DF^{-1} (row-major dim x dim), Jw */
  double *Gk = (double *)malloc(sizeof(double) * E * nd * d * d);
  double *Jw = (double *)malloc(sizeof(double) * E * nd);
  for (long e = 0; e < E; e++) {
    long cv[8];
    elem_corners(dim, nx, e, cv);
    for (int q = 0; q < nd; q++) {
      int loc[3] = {q % n, (q / n) % n, q / (n * n)};
      double xi[3], N8[8], dN[24], DF[9], Di[9], wq = 1.0;
      for (int a = 0; a < dim; a++) { xi[a] = xg[loc[a]]; wq *= wg[loc[a]]; }
      multilinear(dim, xi, N8, dN);
      for (int a = 0; a < dim; a++)
        for (int c = 0; c < dim; c++) {
          double s = 0.0;
          for (int k = 0; k < (1 << dim); k++) s += dN[k * dim + c] * vert[cv[k] * dim + a];
          DF[a * dim + c] = s;
        }
      double J = det_inv(dim, DF, Di);
      long o = e * nd + q;
      Jw[o] = J * wq;
      for (int a = 0; a < d * d; a++) Gk[o * d * d + a] = wq * J * Di[a];
    }
  }
  /* assembled mass and damping blocks per node: D[c][g] for copy c (0 = u),
   * S[g][(c1*d + c2)] */
  double *Dn = (double *)calloc((size_t)d * N, sizeof(double));
  double *Sn = (double *)calloc((size_t)N * d * d, sizeof(double));
  unsigned char *inpml = (unsigned char *)calloc((size_t)N, 1);
  for (long e = 0; e < E; e++)
    for (int q = 0; q < nd; q++) {
      long o = e * nd + q;
      int g = elem_nodes[o];
      double m = rho[e] * Jw[o];
      Dn[g] += m;
      double *S = Sn + (long)g * d * d;
      if (!pml[e]) {
        if (sigma) S[0] += sigma[e] * Jw[o];
        continue;
      }
      inpml[g] = 1;
      for (int c = 1; c < d; c++) Dn[(long)c * N + g] += m;
      const double *t = tau + o * d;
      if (dim == 2) {
        double sp = rho[e] * (0.5 * (t[0] + t[1])) * Jw[o];
        double sm = rho[e] * (0.5 * (t[0] - t[1])) * Jw[o];
        S[0] += sp; S[3] += sp; S[1] += sm; S[2] += sm;
      } else {
        double s0 = rho[e] * (t[0] + t[1] + t[2]) * Jw[o] / 3.0;
        double d1 = rho[e] * (t[0] - t[1]) * Jw[o] / 3.0;
        double d2 = rho[e] * (t[0] - t[2]) * Jw[o] / 3.0;
        double d3 = rho[e] * (t[1] - t[2]) * Jw[o] / 3.0;
        S[0] += s0; S[1] += d1 - d3; S[2] += d2 + d3;
        S[3] += d1; S[4] += s0 + d3; S[5] += d1;
        S[6] += d2; S[7] += d2; S[8] += s0 - d3;
      }
    }
  for (long g = 0; g < N; g++)
    if (dirichlet[g]) {
      Dn[g] = 1.0;
      for (int c = 0; c < d; c++) { Sn[g * d * d + c] = 0.0; Sn[g * d * d + c * d] = 0.0; }
    }
  double *U = (double *)calloc((size_t)d * N, sizeof(double));
  double *Pr = (double *)calloc((size_t)d * N, sizeof(double));
  double *V = (double *)calloc((size_t)E * nd * d, sizeof(double));
  for (int it = 0; it <= nsteps; it++) {
    for (int s = 0; s < nrec; s++) {
      double acc = 0.0;
      long e = rec_elem[s];
      for (int q = 0; q < nd; q++) acc += rec_w[(long)s * nd + q] * U[elem_nodes[e * nd + q]];
      traces[(long)it * nrec + s] = acc;
    }
    if (it == nsteps) break;
    /* U^{n+1}: Pr holds R^S V^{n+1/2} */
    for (long g = 0; g < N; g++) {
      if (b) Pr[g] += f[it] * b[g];
      int k = inpml[g] ? d : 1;
      double A[9], P[9], y[3], x[3];
      for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) {
          double sv = Sn[g * d * d + i * d + j];
          double dg = (i == j) ? Dn[(long)i * N + g] : 0.0;
          A[i * k + j] = dg - 0.5 * dt * sv;
          P[i * k + j] = dg + 0.5 * dt * sv;
        }
      for (int i = 0; i < k; i++) {
        double s = dt * Pr[(long)i * N + g];
        for (int j = 0; j < k; j++) s += A[i * k + j] * U[(long)j * N + g];
        y[i] = s;
      }
      or_block_solve(k, P, y, x);
      for (int i = 0; i < k; i++) U[(long)i * N + g] = x[i];
      if (dirichlet[g]) U[g] = 0.0;
    }
    /* V^{n+3/2} and the next R^S V */
    memset(Pr, 0, sizeof(double) * d * N);
    for (long e = 0; e < E; e++) {
      double ue[OR_MAXN * OR_MAXN * OR_MAXN];
      double fl[3][3 * OR_MAXN * OR_MAXN * OR_MAXN];
      for (int q = 0; q < nd; q++) ue[q] = U[elem_nodes[e * nd + q]];
      for (int q = 0; q < nd; q++) {
        int i = q % n, j = (q / n) % n, kk = q / (n * n);
        double gr[3] = {0, 0, 0};
        for (int m = 0; m < n; m++) {
          gr[0] += Dm[i * n + m] * ue[m + n * j + n * n * kk];
          gr[1] += Dm[j * n + m] * ue[i + n * m + n * n * kk];
          if (dim == 3) gr[2] += Dm[kk * n + m] * ue[i + n * j + n * n * m];
        }
        long o = e * nd + q;
        const double *G = Gk + o * d * d;
        double *v = V + o * d;
        double Bq = Jw[o] / mu[e];
        for (int a = 0; a < d; a++) {
          double rv = 0.0; /* (w J DF^{-T} grad u)_a = sum_c G[c][a] gr_c */
          for (int c = 0; c < d; c++) rv += G[c * d + a] * gr[c];
          double sv = pml[e] ? tau[o * d + a] * Jw[o] / mu[e] : 0.0;
          double pv = dt * rv + (Bq - 0.5 * dt * sv) * v[a];
          v[a] = pv * (1.0 / (Bq + 0.5 * dt * sv));
        }
        /* fluxes -G P v for the copies */
        for (int c = 0; c < (pml[e] ? d : 1); c++) {
          double pvv[3] = {v[0], v[1], dim == 3 ? v[2] : 0.0};
          if (c == 1) { pvv[1] = -v[1]; if (dim == 3) pvv[2] = 0.0; }
          if (c == 2) { pvv[1] = 0.0; pvv[2] = -v[2]; }
          for (int a = 0; a < d; a++) {
            double s = 0.0;
            for (int cc = 0; cc < d; cc++) s += G[a * d + cc] * pvv[cc];
            fl[c][q * d + a] = -s;
          }
        }
      }
      for (int c = 0; c < (pml[e] ? d : 1); c++)
        for (int q = 0; q < nd; q++) {
          int i = q % n, j = (q / n) % n, kk = q / (n * n);
          double s = 0.0;
          for (int m = 0; m < n; m++) {
            s += Dm[m * n + i] * fl[c][(m + n * j + n * n * kk) * d + 0];
            s += Dm[m * n + j] * fl[c][(i + n * m + n * n * kk) * d + 1];
            if (dim == 3) s += Dm[m * n + kk] * fl[c][(i + n * j + n * n * m) * d + 2];
          }
          Pr[(long)c * N + elem_nodes[e * nd + q]] += s;
        }
    }
  }
  memcpy(uN, U, sizeof(double) * N);
  free(Gk); free(Jw); free(Dn); free(Sn); free(inpml); free(U); free(Pr); free(V);
}
