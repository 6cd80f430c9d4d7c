// Internal declarations shared by the host setup (mjw_host.cpp), the kernels
// (mjw_kernels.hip) and the C ABI (mjw_api.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mjw {

constexpr int kMaxOrder = 7;      // Q1..Q7 elements (n = order+1 <= 8)
constexpr int kMaxN = kMaxOrder + 1;

// ----------------------------------------------------------------- host math
// Gauss-Lobatto rule of [0,1] for `order` (order+1 points, weights sum 1).
void gl_rule(int order, double *x, double *w);
// D[a*n+b] = phi_b'(x_a) for the GL Lagrange basis.
void gl_deriv(int order, const double *x, double *D);
// Lagrange basis values phi_i(t) and derivatives on the GL nodes.
void lagrange(int order, const double *x, double t, double *phi, double *dphi);
// Ricker pulse table through the degree-24 GL interpolator.
void ricker_pulse_table(double f0, double t_begin, double dt, int nsteps, double *out);
// out[n] = f^(nb_deriv)(t_begin + n dt + offset), 0 from the cut-off on
void source_table(int kind, double f0, double par, double t_begin, double dt, int nsteps,
                  int nb_deriv, double offset, double *out);
// Stateful Ricker time source: its degree-24 GL interpolant on windows of
// length 0.5/f0 starting at t_init, which move forward with the evaluation
// time; eval(t, k) = k-th derivative of the interpolated primitive (k = 1:
// the pulse), 0 <= k < kNDer.
struct GlTimeSource {
  static constexpr int kR = 24, kNP = kR + 1, kNDer = 10;
  int kind = 1;        // MJW_TIME_* of mjwave.h
  double par = 0.0;    // the kind's parameter (centre time, width, final time, alpha)
  double f0 = 1.0, t0 = 0.0, t1 = -1.0, eps = 0.0;
  bool ready = false;
  double xg[kNP], Dg[kNP * kNP], cinv[kNP];
  double val[kNDer][kNP];
  void init(double freq, double t_init, double eps_time, int kind_ = 1, double par_ = 0.0);
  double primitive(double t) const;
  void generate();
  double eval(double t, int nb_deriv);
};
// the kind's cut-off time (the source is 0 from it on), or NaN for an
// unknown kind or a parameter it cannot take
double time_source_limit(int kind, double f0, double par);

// ----------------------------------------------------------------- host mesh
void mesh_regular_size(int dim, int nx, int order, int64_t *N, int64_t *E);
void mesh_regular(int dim, int nx, int order, double pert, uint64_t seed, double *xyz,
                  int32_t *elem_nodes, uint8_t *dirichlet);

// Greedy first-fit colouring of elements (no two elements of one colour share
// a node).  Returns number of colours or -1.
int color_elements(int64_t E, int nd, int64_t N, const int32_t *elem_nodes,
                   std::vector<int> &color);

// Diagonal operators on host vectors [ncomp][N], parallel over the nodes:
// Y = beta Y + alpha diag(a + cc c) U (a null: coefficient 0; c null: a),
// Y /= a + cc c, and Y = 0 on the nodes with dir set.
void diag_axpby(int64_t N, int ncomp, const double *a, const double *c, double cc, double alpha,
                const double *U, double beta, double *Y);
void diag_solve(int64_t N, int ncomp, const double *a, const double *c, double cc, double *Y);
void diag_mask(int64_t N, int ncomp, const uint8_t *dir, double *Y);

// Point location in an isoparametric Q_r element: Newton on
// x(xi) = sum_q x_q phi_q(xi).  Returns true when the reference point lies in
// [-tol, 1+tol]^dim (xi clamped to [0,1]).
bool locate_point(int dim, int order, const double *xq /*nd*dim*/, const double *p,
                  double *xi);
// tensor basis values phi_q(xi) (nd values)
void tensor_basis(int dim, int order, const double *xi, double *phi);
void tensor_grad(int dim, int order, const double *xq /*nd*dim*/, const double *xi,
                 double *g /*dim*nd*/);

// ------------------------------------------------------------ device launch
enum StateType { kF64 = 0, kF32 = 1 };
enum GeomType { kG64 = 0, kG32 = 1 };
enum Mode { kApply = 0, kLeap = 1 };

// Trilinear (vertex) geometry of a hexahedron: x(xi) = a0 + a1 xi1 + a2 xi2 +
// a3 xi3 + a4 xi1 xi2 + a5 xi2 xi3 + a6 xi1 xi3 + a7 xi1 xi2 xi3 on [0,1]^3.
// Per element (colour-sorted order) kCoef values: a1..a7 (3 components each,
// slots 0..20), mu (slot 21), 0, 0.
constexpr int kCoef = 24;
// Fills coef (E*kCoef, sorted order via perm) and returns true when every
// element's nodes are the trilinear image of the GL grid (within tol times
// the element diameter); false otherwise (coef then undefined).
bool trilinear_coefficients(int order, int64_t E, const int32_t *elem_nodes, const double *xyz,
                            const int64_t *perm, const double *mu_sorted, double tol,
                            std::vector<double> &coef);
enum GeoMode { kGeoStored = 0, kGeoTrilinear = 1 };
// tile codes of the 3-D scalar kernels: EPB > 0 column kernel, -EPB one thread
// per node, kPipeTile + EPB persistent column kernel
constexpr int kPipeTile = 1000;

struct StiffArgs {
  int dim, order, ncomp, epb;         // geometry, elements per block
  int state;                          // StateType
  int geom;                           // GeomType
  int mode;                           // Mode
  int64_t E;                          // plane stride (all elements)
  int64_t e0, e1;                     // element range (one colour)
  const int32_t *map;                 // plane-major [K][E][P], bit31 = first toucher
  const void *G;                      // geometry planes
  const void *Dm;                     // n*n derivative matrix, state precision
  const double *lam, *mu;             // elastic per element (sorted order), FP64
  const void *U;                      // state in
  void *Y;                            // out (apply) / previous->next (leapfrog)
  const void *sdiag;                  // dt^2 / M per node (leapfrog)
  double alpha, beta;                 // apply
  int64_t N;                          // nodes
  int nshots;
  int nxcd;                           // XCD-aware block order (1 = off)
  double *grad;                       // adjoint step: per-element gradient (sorted), or null
  int ilv;                            // several shots interleaved: u[(m*N + g)*S + s]
  int geo;                            // GeoMode (trilinear: 3-D scalar only)
  const void *coef;                   // trilinear: E*kCoef values, geometry precision
  const void *glxw;                   // trilinear: GL points then weights (n each), geometry precision
};
// returns 0 or a HIP error code; `stream` is a hipStream_t
int launch_stiffness(const StiffArgs &a, void *stream);
// 3-D acoustic, several shots in the interleaved layout; StiffArgs::epb =
// shot lanes per column (0: automatic)
int launch_stiffness_shots(const StiffArgs &a, void *stream);
// interleaved acoustic shots with the contractions on the matrix cores (FP32
// state, orders 1..6): tile kTileMfma
constexpr int kTileMfma = 1000;
bool shots_mfma_supported(int order, int state);
int launch_stiffness_shots_mfma(const StiffArgs &a, void *stream);
// supported shot-lane counts, largest first
std::vector<int> tile_choices_shots(int order, int state, int geom);
// isotropic elastic, several shots interleaved
int launch_stiffness_shots_elastic(const StiffArgs &a, void *stream);
std::vector<int> tile_choices_shots_elastic(int order, int state, int geom);
// list of supported elements-per-block for (dim, order, ncomp)
std::vector<int> tile_choices(int dim, int order, int ncomp);
int default_tile(int dim, int order, int ncomp, int state);

// FP64 geometry: acoustic tensor or elastic sqrt(Jw) DF^{-1} planes, Mloc[E*nd].
// rho_q / mu_q (E*nd, sorted element order; may be null) replace the
// per-element rho / mu by values at every quadrature point
int launch_geometry_full(int dim, int order, int ncomp, int64_t E, const int32_t *map,
                         const double *xyz, const double *D, const double *w, const double *rho,
                         const double *mu, const double *rho_q, const double *mu_q, double *G64,
                         float *G32, double *Mloc, int *bad, void *stream);
// state at probe points (FP64 output, [nshots][npts][ncomp])
int launch_probe(int state, const void *U, int ilv, int nshots, int ncomp, int64_t N, int npts,
                 int nd, const int32_t *gid, const double *w, double *out, void *stream);
// mass assembly of one colour range: M[g] (=|+=) Mloc
int launch_mass_assemble(const int32_t *map, const double *Mloc, int64_t E, int dim, int order,
                         int64_t e0, int64_t e1, double *M, void *stream);

struct TailArgs {
  int state;
  int64_t *ctl;                  // [0] = step, [1] = base
  const double *pulse;           // table
  int64_t npulse;
  int nsrc;                      // entries
  const int64_t *src_dof;        // dof (m*N+g)
  const int32_t *src_shot;
  const double *src_val;
  const void *sdiag;             // state precision, per node
  int nrec, nd;
  const int32_t *rec_gid;        // nrec*nd global node ids
  const double *rec_w;           // nrec*nd
  const int32_t *rec_comp;
  const void *U;                 // U^n
  void *P;                       // U^{n+1} being formed (null: no injection)
  double *traces;                // [step-base][S][nrec] (null: none)
  int64_t N;
  int nshots;
  int advance;                   // +1 / -1: step change after, 0: none
  int row_offset;                // trace row = step - base + row_offset
  // adjoint injection: U^{n+1}[node] += s[node] * sum_k w_k g[step][rec_k]
  // (receiver transpose, CSR over unique nodes); applied to shot adj_shot
  int nadj;                      // unique nodes
  const int32_t *adj_node;
  const int32_t *adj_ptr;        // nadj+1
  const int32_t *adj_rec;
  const double *adj_w;
  const double *adj_g;           // [row][adj_nrec]
  int64_t adj_row;               // row of adj_g = step + adj_row
  int adj_nrec;
  int adj_shot;
  int ilv;                       // state layout: shots interleaved (1) or consecutive (0)
};
int launch_tail_n(const TailArgs &a, int ncomp, void *stream);

// elementwise helpers (state precision); vectors are contiguous
int launch_axpby(int state, int64_t n, double a, const void *x, double b, void *y, void *stream);
int launch_convert_to_f64(int state, int64_t n, const void *x, double *y, void *stream);
int launch_convert_from_f64(int state, int64_t n, const double *x, void *y, void *stream);
// caller's layout [S][ncomp][N] (caller's node numbering) <-> solver layout
// (solver node order; shots consecutive, or interleaved [ncomp][N][S] when
// ilv): to_solver gathers through ninv (solver -> caller node), from_solver
// through nperm (caller -> solver node); x and y distinct
int launch_to_solver(int state, int S, int ncomp, int64_t N, int ilv, const int32_t *ninv,
                     const void *x, void *y, void *stream);
int launch_from_solver(int state, int S, int ncomp, int64_t N, int ilv, const int32_t *nperm,
                       const void *x, void *y, void *stream);
// one FP64 vector in the caller's numbering ([ncomp][N]) <-> shot s of the
// solver layout (S shots, interleaved when ilv) in the state precision
int launch_to_solver_f64(int state, int S, int s, int ncomp, int64_t N, int ilv,
                         const int32_t *ninv, const double *x, void *y, void *stream);
int launch_from_solver_f64(int state, int S, int s, int ncomp, int64_t N, int ilv,
                           const int32_t *nperm, const void *x, double *y, void *stream);
// y[i] *= c*d[node(i)], node(i) = (ilv ? i / S : i) % N
int launch_scale_by(int state, int64_t n, int S, int ilv, int64_t N, const double *d, double c,
                    void *y, void *stream);
// y[i] = 0 where dir[node(i)]
// node-wise steps of the fourth-order modified-equation leapfrog (k_me4)
enum MeOp { kMeVel = 0, kMeSolve = 1, kMeSub = 2, kMeNew = 3, kMeTaylor = 4, kMePre = 5 };
int launch_me4(int state, int op, int64_t n, int S, int ilv, int64_t N, const double *D,
               const double *Sd, const uint8_t *dir, const void *a, const void *b, const void *x,
               const void *u, const void *w, double c0, double c1, double c2, void *y,
               void *stream);
int launch_mask(int state, int64_t n, int S, int ilv, int64_t N, const uint8_t *dir, void *y,
                void *stream);
int launch_make_sdiag(int state, int64_t N, const double *M, const uint8_t *dir, double dt2,
                      void *s, void *stream);
// ctl[0] = step, ctl[1] = base on the device (a kernel, no host copy)
int launch_set_ctl(int64_t *ctl, int64_t step, int64_t base, void *stream);
// damped nodes: mode 0 save = y, 1 y += coef (save - u), 2 y += coef u
int launch_damp(int state, int mode, int64_t ndamp, int ncomp, int S, int ilv, int64_t N,
                const int32_t *idx, const double *coef, void *y, void *save, const void *u,
                void *stream);
// per shot: out[s] = sum_i w[i%N] x_i y_i  /  sum_i w[i%N] (a_i-b_i)^2, FP64,
// deterministic; part = scratch of reduction_scratch_doubles(S) doubles
int launch_dot_part(int state, int64_t N, int ncomp, int S, int ilv, const double *w,
                    const void *x, const void *y, double *out, double *part, void *stream);
int launch_dot_diff_part(int state, int64_t N, int ncomp, int S, int ilv, const double *w,
                         const void *a, const void *b, double *out, double *part, void *stream);
int reduction_scratch_doubles(int S);
// G planes *= sc[e] (per element, sorted order); g32: float planes
int launch_scale_planes(int g32, void *G, int64_t nplanes, int64_t E, int P, const double *sc,
                        void *stream);

// ------------------------------------------------- PML layers (LEAP_FROG_PML)
struct PmlArgs {
  int dim, order, state;
  int64_t E, e0, e1, N;                // elements (plane stride), colour range, nodes
  const int32_t *map;                  // element map (plane-major, bit 31 ignored)
  const double *G;                     // [d*d][E*nd] w J DF^{-1}, row-major
  const double *B;                     // [E*nd] J w / mu
  const double *tau;                   // [d][E*nd] PML damping (PML elements)
  const uint8_t *pml;                  // [E] element inside the layers (sorted order)
  const double *Dm;                    // n*n GL derivative matrix
  void *U, *V, *Pr;                    // [d][N], [d][E*nd], [d][N] (state precision)
  double dt;
  int64_t npn;                         // nodes of PML elements
  const int32_t *pnode;                // [npn] their solver indices
  const double *A, *Pinv;              // [npn][d][d] D - dt/2 S, (D + dt/2 S)^{-1}
  const uint8_t *dir, *blk;            // [N] Dirichlet node, PML node
  const double *sa, *sinv;             // [N] D - dt/2 S, 1/(D + dt/2 S) (0 on Dirichlet)
};
int launch_pml_elem(const PmlArgs &a, void *stream);
int launch_pml_nodes(const PmlArgs &a, void *stream);

// ---------------------------------------------------------------- batching
struct BatchProb {          // device descriptor of one problem of a batch
  const int32_t *map;
  const void *G, *Dm, *sdiag;
  void *buf[2];             // the problem's two state buffers
  const void *coef, *glxw;  // trilinear geometry (geo == kGeoTrilinear)
  int64_t E, N;
  int order, geo;
};
struct BatchItem {          // one block of work: elements [eb, min(eb+EPB, e1))
  int32_t prob, pad;
  int64_t eb, e1;
};
int batch_epb(int order);
int launch_batch3d(int state, int geom, const BatchProb *probs, const BatchItem *items,
                   int nitems, int parity, void *stream);
// grouped launches: one per (colour, order), tile 0 / 1 / 2 = blocks of about
// 256 / 512 / 128 threads
int batch_group_epb(int order, int tile);
int launch_batch3d_order(int state, int geom, int order, int tile, const BatchProb *probs,
                         const BatchItem *items, int nitems, int parity, void *stream);
// one tail block per problem: args[blockIdx.x]
int launch_tail_batch(int state, const TailArgs *args, int nprob, void *stream);

}  // namespace mjw
