// C ABI of the wave solver (include/mjwave.h): problem setup, operators,
// sources / receivers, the fused leapfrog step loop (optionally captured in a
// HIP graph), energy and CFL.  One problem = one device + one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <array>
#include <map>
#include <mutex>
#include <numeric>
#include <unordered_set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mjwave.h"
#include "mjw_internal.hpp"

using namespace mjw;

namespace {
thread_local std::string g_err;
}

struct mjw_problem {
  int device = 0;
  hipStream_t stream = nullptr;
  int dim = 3, order = 1, n = 2, ncomp = 1, nshots = 1, prec = MJW_FP64;
  int state = kF64, geom = kG64;
  int64_t N = 0, E = 0;
  int nd = 0, P = 0;
  int ncolor = 0;
  std::vector<int64_t> coff;    // colour offsets in sorted order
  std::vector<int64_t> perm;    // sorted -> original element
  // device vectors use the solver node order: nperm[g] = solver index of
  // node g of the caller's numbering, ninv = its inverse
  std::vector<int32_t> nperm, ninv;
  // several shots of a 3-D problem: interleaved state u[(m*N + g)*S + s]
  int ilv = 0;
  int st = 0;                   // shot lanes of the interleaved kernel (0: automatic)
  int32_t *d_nperm = nullptr, *d_ninv = nullptr;
  int epb = 0;
  int nxcd = 8;                 // MI355X: 8 XCDs, round-robin block dispatch
  // host copies (original element order)
  std::vector<double> xyz;
  std::vector<int32_t> en;
  std::vector<uint8_t> dir;
  std::vector<double> bbox;     // E * 2 * dim
  // point location: uniform buckets over the mesh box, each listing (in
  // element order) the elements whose box meets it; built on first use
  std::vector<int32_t> loc_ptr, loc_elem;
  int loc_n[3] = {0, 0, 0};
  double loc_lo[3] = {0, 0, 0}, loc_h[3] = {1, 1, 1};
  // probes: points where mjw_probe samples the current iterate
  int nprobe = 0;
  int32_t *d_pr_gid = nullptr;
  double *d_pr_w = nullptr, *d_pr_out = nullptr;
  std::vector<double> mass;     // N, Dirichlet rows 1
  // device
  int32_t *d_map = nullptr;
  void *d_G = nullptr;
  // trilinear geometry (3-D scalar problems whose elements are straight-sided)
  bool tri_ok = false;
  int geo = kGeoStored;
  bool nodal_media = false;           // rho, mu given per quadrature point
  // damping S_h = sigma J w per node (mjw_set_damping), solver order
  std::vector<double> damp_S;
  std::vector<double> damp_c;         // the same in the caller's numbering, 0 on Dirichlet rows
  int64_t ndamp = 0;                  // damped, non-Dirichlet nodes
  double *d_Meff = nullptr, *d_damp_k = nullptr;
  int32_t *d_damp_idx = nullptr;
  void *d_damp_save = nullptr;
  std::vector<double> hcoef;          // E*kCoef, sorted order, FP64
  void *d_coef = nullptr, *d_glxw = nullptr;
  double *d_M = nullptr, *d_Minv = nullptr, *d_isqM = nullptr;
  void *d_sdiag = nullptr;
  uint8_t *d_dir = nullptr;
  double *d_lam = nullptr, *d_mu = nullptr, *d_rho = nullptr;
  void *d_D = nullptr;
  void *d_U[2] = {nullptr, nullptr};
  int cur = 0;
  // fourth-order modified equation (mjw_set_time_order): U^{n-1}, work
  // vectors V, X, Z, the source tables dt^2 f, dt^2 f', dt^2 f'' and S per node
  int torder = 2;
  void *d_U3 = nullptr;
  void *d_me[3] = {nullptr, nullptr, nullptr};
  void *d_sdiag4 = nullptr;           // dt^2 / (12 D), 0 on Dirichlet nodes
  double *d_pulse_me = nullptr;
  double *d_Sme = nullptr;
  int me_phase = 0;                   // steps mod 6: the cycle of the three levels
  std::vector<double> h_pulse[3];     // f, f', f'' at t_n + dt (host copies)
  void *d_tmp = nullptr, *d_tmp2 = nullptr;
  double *d_stage = nullptr;          // host-vector staging (one FP64 vector)
  double *d_part = nullptr, *d_red = nullptr;
  // leapfrog
  double dt = 0.0, t_begin = 0.0;
  bool inited = false;
  int64_t *d_ctl = nullptr;
  int64_t step = 0;
  double *d_pulse = nullptr;
  int64_t npulse = 0;
  int nsrc = 0;
  // host copy of the sparse source vectors (caller's numbering m*N + g)
  std::vector<int64_t> h_src_dof;
  std::vector<int32_t> h_src_shot;
  std::vector<double> h_src_val;
  // time function of the operator-level source entry (mjw_set_time_source)
  GlTimeSource tsrc;
  bool has_tsrc = false;
  double tlimit = 0.0;
  int64_t *d_src_dof = nullptr;
  int32_t *d_src_shot = nullptr;
  double *d_src_val = nullptr;
  int nrec = 0;
  int32_t *d_rec_gid = nullptr, *d_rec_comp = nullptr;
  double *d_rec_w = nullptr;
  double *d_traces = nullptr;
  int64_t traces_cap = 0;
  // graph of an even number of steps, valid for one parity / traces buffer
  hipGraphExec_t gexec = nullptr;
  int g_steps = 0, g_parity = -1;
  double *g_traces = nullptr;
  bool g_with_traces = false;
  double ms_step = 0.0, ms_stiff = 0.0;
  size_t bytes = 0;
  std::string err;
  // adjoint
  bool zero_ic = true;
  std::vector<void *> snaps;          // pairs (U^k, U^{k+1}), 2*N*ncomp each
  std::vector<int64_t> snap_step;     // k of each pair
  int64_t adj_nsteps = 0;
  bool adj_ready = false;
  void *d_pair[2] = {nullptr, nullptr};
  double *d_grad = nullptr, *d_adjg = nullptr;
  int64_t adjg_cap = 0;
  int nadj = 0;                       // receiver transpose (CSR over unique nodes)
  int32_t *d_adj_node = nullptr, *d_adj_ptr = nullptr, *d_adj_rec = nullptr;
  double *d_adj_w = nullptr;
  hipGraphExec_t bexec = nullptr;     // two backward steps
  double ms_adj_fwd = 0.0, ms_adj_bwd = 0.0;
  // PML layers: first-order split system (mjw_set_pml)
  bool pml = false;
  std::vector<double> pml_D;          // [d][N] lumped mass of u, u*, u** (solver order)
  std::vector<double> pml_S;          // [npn][d][d] damping blocks of the PML nodes
  std::vector<int32_t> pml_node;      // [npn] their solver indices
  double *d_pml_G = nullptr, *d_pml_B = nullptr, *d_pml_tau = nullptr, *d_pml_Dm = nullptr;
  uint8_t *d_pml_flag = nullptr, *d_pml_blk = nullptr;
  void *d_pml_V = nullptr, *d_pml_Pr = nullptr;
  int32_t *d_pml_node = nullptr;
  double *d_pml_A = nullptr, *d_pml_Pinv = nullptr, *d_pml_sa = nullptr, *d_pml_sinv = nullptr;
};

#define FAIL(p, code, msg)                  \
  do {                                      \
    std::string _m = (msg);                 \
    if (p) (p)->err = _m; else g_err = _m;  \
    return (code);                          \
  } while (0)

#define HIPCHK(p, x)                                                                 \
  do {                                                                               \
    hipError_t _e = (x);                                                             \
    if (_e != hipSuccess)                                                            \
      FAIL(p, MJW_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(_e) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));           \
  } while (0)

#define KCHK(p, x)                                                                   \
  do {                                                                               \
    int _e = (x);                                                                    \
    if (_e != 0)                                                                     \
      FAIL(p, MJW_ERR_HIP, std::string("kernel launch failed (") + std::to_string(_e) + \
                               ") at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

// ------------------------------------------------ step loop without allocation
// mjw_leapfrog_advance_static stages the trace rows in `scratch`, a file-scope
// device array of kStaticTraceDoubles doubles (8 MiB) that every GPU the
// library runs on holds until the process ends.  Host code never uses the
// array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device with that device current (in
// mjw_create), and is kept per device in g_scratch.
constexpr int64_t kStaticTraceDoubles = int64_t(1) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kStaticTraceDoubles];
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices] = {};

// the calling thread's current device must be `device`
static hipError_t setup_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch[device] = static_cast<double *>(ptr);
  return e;
}

static double *scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

static size_t ssize(const mjw_problem *p) { return p->state == kF64 ? 8 : 4; }
static int64_t nstate(const mjw_problem *p) { return (int64_t)p->nshots * p->ncomp * p->N; }

template <typename T>
static hipError_t dalloc(mjw_problem *p, T **ptr, size_t bytes) {
  hipError_t e = hipMalloc((void **)ptr, bytes ? bytes : 8);
  if (e == hipSuccess) p->bytes += bytes;
  return e;
}

static void dfree(void *ptr) {
  if (ptr) (void)hipFree(ptr);
}

extern "C" int mjw_version(void) { return 10000; }

extern "C" const char *mjw_last_error(const mjw_problem *p) {
  return p ? p->err.c_str() : g_err.c_str();
}

// ------------------------------------------------------------------ meshes
extern "C" int mjw_mesh_regular_size(int dim, int nx, int order, int64_t *n_nodes,
                                     int64_t *n_elem) {
  if ((dim != 2 && dim != 3) || nx < 1 || order < 1 || order > kMaxOrder || !n_nodes || !n_elem)
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_mesh_regular_size: bad arguments");
  mesh_regular_size(dim, nx, order, n_nodes, n_elem);
  if (*n_nodes >= (int64_t)1 << 31)
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mesh too large for 31-bit node ids");
  return MJW_OK;
}

extern "C" int mjw_mesh_regular(int dim, int nx, int order, double pert, uint64_t seed,
                                double *xyz, int32_t *elem_nodes, uint8_t *dirichlet) {
  int64_t N, E;
  int rc = mjw_mesh_regular_size(dim, nx, order, &N, &E);
  if (rc) return rc;
  if (!xyz || !elem_nodes || !dirichlet || !(pert >= 0.0) || pert >= 0.5)
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_mesh_regular: bad arguments");
  mesh_regular(dim, nx, order, pert, seed, xyz, elem_nodes, dirichlet);
  return MJW_OK;
}

extern "C" int mjw_gauss_lobatto(int order, double *points, double *weights) {
  if (order < 1 || order > 24 || !points || !weights)
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_gauss_lobatto: order in 1..24, non-NULL arrays");
  gl_rule(order, points, weights);
  return MJW_OK;
}

extern "C" int mjw_ricker_pulse_table(double f0, double t_begin, double dt, int nsteps,
                                      double *out) {
  if (!(f0 > 0) || !(dt > 0) || nsteps < 0 || (!out && nsteps > 0))
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_ricker_pulse_table: bad arguments");
  ricker_pulse_table(f0, t_begin, dt, nsteps, out);
  return MJW_OK;
}

extern "C" int mjw_temporal_source_table(int kind, double f0, double param, double t_begin,
                                         double dt, int nsteps, int nb_deriv, double offset,
                                         double *out) {
  if (!(f0 > 0) || !(dt > 0) || nsteps < 0 || (!out && nsteps > 0) || nb_deriv < 0 ||
      nb_deriv >= GlTimeSource::kNDer || !std::isfinite(offset) || !std::isfinite(t_begin))
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_temporal_source_table: bad arguments");
  if (std::isnan(time_source_limit(kind, f0, param)))
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG,
         "mjw_temporal_source_table: unknown kind or a parameter it cannot take");
  source_table(kind, f0, param, t_begin, dt, nsteps, nb_deriv, offset, out);
  return MJW_OK;
}

extern "C" int mjw_time_source_table(double f0, double t_begin, double dt, int nsteps,
                                     int nb_deriv, double offset, double *out) {
  return mjw_temporal_source_table(MJW_TIME_RICKER, f0, 0.0, t_begin, dt, nsteps, nb_deriv,
                                   offset, out);
}

// --------------------------------------------------------------- lifecycle
static void release(mjw_problem *p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  dfree(p->d_U3);
  dfree(p->d_sdiag4);
  for (void *&v : p->d_me) dfree(v);
  dfree(p->d_pulse_me);
  dfree(p->d_Sme);
  dfree(p->d_coef);
  dfree(p->d_glxw);
  dfree(p->d_stage);
  dfree(p->d_nperm);
  dfree(p->d_ninv);
  void *ptrs[] = {p->d_map, p->d_G, p->d_M, p->d_Minv, p->d_isqM, p->d_sdiag, p->d_dir,
                  p->d_lam, p->d_mu, p->d_rho, p->d_D, p->d_U[0], p->d_U[1], p->d_tmp,
                  p->d_tmp2, p->d_part, p->d_red, p->d_ctl, p->d_pulse, p->d_src_dof,
                  p->d_src_shot, p->d_src_val, p->d_rec_gid, p->d_rec_comp, p->d_rec_w,
                  p->d_traces, p->d_pr_gid, p->d_pr_w, p->d_pr_out, p->d_Meff, p->d_damp_k,
                  p->d_damp_idx, p->d_damp_save};
  for (void *q : ptrs) dfree(q);
  for (void *q : p->snaps) dfree(q);
  void *aptrs[] = {p->d_pair[0], p->d_pair[1], p->d_grad, p->d_adjg, p->d_adj_node,
                   p->d_adj_ptr, p->d_adj_rec, p->d_adj_w};
  for (void *q : aptrs) dfree(q);
  if (p->bexec) (void)hipGraphExecDestroy(p->bexec);
  void *pptrs[] = {p->d_pml_G, p->d_pml_B, p->d_pml_tau, p->d_pml_Dm, p->d_pml_flag,
                   p->d_pml_blk, p->d_pml_V, p->d_pml_Pr, p->d_pml_node, p->d_pml_A,
                   p->d_pml_Pinv, p->d_pml_sa, p->d_pml_sinv};
  for (void *q : pptrs) dfree(q);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

extern "C" int mjw_destroy(mjw_problem *p) {
  release(p);
  return MJW_OK;
}

// coefficient table and GL points / weights in the geometry precision
static hipError_t upload_trilinear(mjw_problem *p, const double *xg, const double *wg) {
  const size_t nc = p->hcoef.size();
  hipError_t e = hipSuccess;
  if (p->geom == kG64) {
    e = hipMemcpy(p->d_coef, p->hcoef.data(), nc * 8, hipMemcpyHostToDevice);
  } else {
    std::vector<float> f(p->hcoef.begin(), p->hcoef.end());
    e = hipMemcpy(p->d_coef, f.data(), nc * 4, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess || !xg) return e;
  double gw[2 * kMaxN];
  for (int i = 0; i < p->n; ++i) { gw[i] = xg[i]; gw[p->n + i] = wg[i]; }
  if (p->geom == kG64) return hipMemcpy(p->d_glxw, gw, 2 * p->n * 8, hipMemcpyHostToDevice);
  float gf[2 * kMaxN];
  for (int i = 0; i < 2 * p->n; ++i) gf[i] = (float)gw[i];
  return hipMemcpy(p->d_glxw, gf, 2 * p->n * 4, hipMemcpyHostToDevice);
}

// p->mass holds the assembled mass in solver order (as downloaded): convert
// it to the caller's order, set Dirichlet rows to 1 (TimeAcoustic.cxx:936-937)
// and upload M, 1/M and 1/sqrt(M) (0 on Dirichlet rows) in solver order.
static hipError_t finish_mass(mjw_problem *p) {
  const int64_t n_nodes = p->N;
  {
    std::vector<double> ms_solver(p->mass);
    for (int64_t g = 0; g < n_nodes; ++g) p->mass[g] = ms_solver[p->nperm[g]];
  }
  std::vector<double> msv(n_nodes), minv(n_nodes), isq(n_nodes);
  for (int64_t g = 0; g < n_nodes; ++g) {
    if (p->dir[g]) p->mass[g] = 1.0;
    const int64_t i = p->nperm[g];
    msv[i] = p->mass[g];
    minv[i] = 1.0 / p->mass[g];
    isq[i] = p->dir[g] ? 0.0 : 1.0 / std::sqrt(p->mass[g]);
  }
  hipError_t e = hipMemcpyAsync(p->d_M, msv.data(), n_nodes * 8, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->d_Minv, minv.data(), n_nodes * 8, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(p->d_isqM, isq.data(), n_nodes * 8, hipMemcpyHostToDevice, p->stream);
  // the host vectors must outlive the copies
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  return e;
}

extern "C" int mjw_create(mjw_problem **out, int device, int dim, int order, int ncomp,
                          int64_t n_nodes, const double *node_xyz, int64_t n_elem,
                          const int32_t *elem_nodes, const double *rho, const double *mu,
                          const double *lambda, const uint8_t *dirichlet, int nshots,
                          int precision) {
  if (!out) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_create: out is NULL");
  *out = nullptr;
  if (dim != 2 && dim != 3) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "dim must be 2 or 3");
  if (order < 1 || order > kMaxOrder)
    FAIL((mjw_problem *)nullptr, MJW_ERR_UNSUPPORTED, "order must be in 1..7");
  if (ncomp != 1 && ncomp != dim)
    FAIL((mjw_problem *)nullptr, MJW_ERR_UNSUPPORTED,
         "ncomp must be 1 (acoustic) or the dimension (elastic)");
  if (n_nodes <= 0 || n_nodes >= ((int64_t)1 << 31) || n_elem <= 0 || !node_xyz || !elem_nodes ||
      !rho || !mu || (ncomp > 1 && !lambda))
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_create: bad sizes or NULL arrays");
  if (nshots < 1 || nshots > 4096) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "bad nshots");
  if (precision != MJW_FP64 && precision != MJW_FP32 && precision != MJW_FP32_G64)
    FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "bad precision");
  {
    // buffer resources of the kernels address at most 2 GiB each (checked
    // before any array is read)
    const int64_t P = (int64_t)(order + 1) * (order + 1);
    const int64_t K = dim == 2 ? 1 : order + 1;
    const int64_t sb = precision == MJW_FP64 ? 8 : 4, gb = precision == MJW_FP32 ? 4 : 8;
    const int64_t ngc = ncomp > 1 ? dim * dim : (dim == 2 ? 3 : 6);
    const int64_t lim = ((int64_t)1 << 31) - 4096;
    // several 3-D shots are interleaved: one array of S*ncomp*N values
    const int64_t sv = (dim == 3 && nshots > 1) ? nshots : 1;
    if (sv * n_nodes * ncomp * sb >= lim || n_elem * P * K * 4 >= lim ||
        n_elem * P * ngc * gb >= lim)
      FAIL((mjw_problem *)nullptr, MJW_ERR_UNSUPPORTED,
           "problem too large for one device problem (2 GiB extent per array); split it into "
           "several problems");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    FAIL((mjw_problem *)nullptr, MJW_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "bad device");

  mjw_problem *p = new mjw_problem();
  p->device = device;
  p->dim = dim;
  p->order = order;
  p->n = order + 1;
  p->ncomp = ncomp;
  p->nshots = nshots;
  p->ilv = (dim == 3 && nshots > 1) ? 1 : 0;
  p->prec = precision;
  p->state = precision == MJW_FP64 ? kF64 : kF32;
  p->geom = precision == MJW_FP32 ? kG32 : kG64;
  p->N = n_nodes;
  p->E = n_elem;
  p->P = p->n * p->n;
  p->nd = dim == 2 ? p->P : p->P * p->n;
  const int nd = p->nd;
  auto fail = [&](int code, const std::string &m) {
    g_err = m;
    release(p);
    return code;
  };
  for (int64_t i = 0; i < n_elem * nd; ++i)
    if (elem_nodes[i] < 0 || elem_nodes[i] >= n_nodes)
      return fail(MJW_ERR_ARG, "elem_nodes entry out of range");
  for (int64_t e = 0; e < n_elem; ++e)
    if (!(rho[e] > 0) || !std::isfinite(mu[e]) || !(mu[e] > 0) ||
        (ncomp > 1 && !std::isfinite(lambda[e])))
      return fail(MJW_ERR_ARG, "media must be finite with rho, mu > 0");

  if (hipSetDevice(device) != hipSuccess) return fail(MJW_ERR_HIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(MJW_ERR_HIP, "hipStreamCreate failed");
  if (setup_scratch(device) != hipSuccess)
    return fail(MJW_ERR_HIP, "hipGetSymbolAddress of the static trace buffer failed");

  // host copies
  p->xyz.assign(node_xyz, node_xyz + n_nodes * dim);
  p->en.assign(elem_nodes, elem_nodes + n_elem * nd);
  p->dir.assign(n_nodes, 0);
  if (dirichlet)
    for (int64_t g = 0; g < n_nodes; ++g) p->dir[g] = dirichlet[g] ? 1 : 0;
  p->bbox.resize(n_elem * 2 * dim);
  for (int64_t e = 0; e < n_elem; ++e) {
    double *b = &p->bbox[e * 2 * dim];
    for (int a = 0; a < dim; ++a) { b[a] = 1e300; b[dim + a] = -1e300; }
    for (int q = 0; q < nd; ++q) {
      const double *x = &p->xyz[(int64_t)p->en[e * nd + q] * dim];
      for (int a = 0; a < dim; ++a) {
        b[a] = std::min(b[a], x[a]);
        b[dim + a] = std::max(b[dim + a], x[a]);
      }
    }
  }

  // colouring, colour-sorted order, first-toucher flags, plane-major map
  std::vector<int> color;
  const int nc = color_elements(n_elem, nd, n_nodes, elem_nodes, color);
  if (nc < 0) return fail(MJW_ERR_ARG, "colouring failed (more than 64 colours)");
  p->ncolor = nc;
  p->coff.assign(nc + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e) p->coff[color[e] + 1]++;
  for (int c = 0; c < nc; ++c) p->coff[c + 1] += p->coff[c];
  p->perm.resize(n_elem);
  {
    std::vector<int64_t> pos(p->coff.begin(), p->coff.end() - 1);
    for (int64_t e = 0; e < n_elem; ++e) p->perm[pos[color[e]]++] = e;
  }
  // solver node order: nodes grouped by the set of colours whose elements
  // touch them (each colour pass then reads whole groups, so a cache line
  // carries only nodes of the current pass), inside a group by first touch
  // in the kernels' element order
  {
    std::vector<uint64_t> mask(n_nodes, 0);
    for (int64_t e = 0; e < n_elem; ++e)
      for (int q = 0; q < nd; ++q) mask[elem_nodes[e * nd + q]] |= 1ull << color[e];
    std::vector<int32_t> by_rank;
    by_rank.reserve(n_nodes);
    std::vector<uint8_t> seen(n_nodes, 0);
    for (int64_t es = 0; es < n_elem; ++es) {
      const int64_t e = p->perm[es];
      for (int q = 0; q < nd; ++q) {
        const int32_t g = elem_nodes[e * nd + q];
        if (!seen[g]) { seen[g] = 1; by_rank.push_back(g); }
      }
    }
    for (int64_t g = 0; g < n_nodes; ++g)
      if (!seen[g]) by_rank.push_back((int32_t)g);  // nodes of no element
    // the distinct colour sets (a few dozen), sorted; one pass over the nodes
    std::vector<uint64_t> masks;
    {
      std::unordered_set<uint64_t> distinct;
      uint64_t last = ~0ull;
      for (uint64_t m : mask)
        if (m != last) {
          last = m;
          if (distinct.insert(m).second) masks.push_back(m);
        }
      std::sort(masks.begin(), masks.end());
    }
    std::vector<int64_t> off(masks.size() + 1, 0);
    std::vector<int32_t> cid(n_nodes);
    for (int64_t g = 0; g < n_nodes; ++g) {
      cid[g] = (int32_t)(std::lower_bound(masks.begin(), masks.end(), mask[g]) - masks.begin());
      off[cid[g] + 1]++;
    }
    for (size_t c = 0; c < masks.size(); ++c) off[c + 1] += off[c];
    p->nperm.resize(n_nodes);
    p->ninv.resize(n_nodes);
    for (int32_t g : by_rank) {
      const int64_t i = off[cid[g]]++;
      p->nperm[g] = (int32_t)i;
      p->ninv[i] = g;
    }
  }
  std::vector<int32_t> hmap((size_t)n_elem * nd);
  {
    std::vector<uint8_t> seen(n_nodes, 0);
    const int K = dim == 2 ? 1 : p->n;
    for (int64_t es = 0; es < n_elem; ++es) {
      const int64_t e = p->perm[es];
      for (int k = 0; k < K; ++k)
        for (int ij = 0; ij < p->P; ++ij) {
          const int q = ij + p->P * k;
          const int32_t g = elem_nodes[e * nd + q];
          uint32_t v = (uint32_t)p->nperm[g];
          if (!seen[g]) { seen[g] = 1; v |= 0x80000000u; }
          hmap[((size_t)k * n_elem + es) * p->P + ij] = (int32_t)v;
        }
    }
  }
  std::vector<double> rs(n_elem), ms(n_elem), ls(ncomp > 1 ? n_elem : 1);
  for (int64_t es = 0; es < n_elem; ++es) {
    rs[es] = rho[p->perm[es]];
    ms[es] = mu[p->perm[es]];
    if (ncomp > 1) ls[es] = lambda[p->perm[es]];
  }
  // 1-D rule
  double xg[kMaxN], wg[kMaxN], Dg[kMaxN * kMaxN];
  gl_rule(order, xg, wg);
  gl_deriv(order, xg, Dg);
  std::vector<float> Df(p->n * p->n);
  for (int i = 0; i < p->n * p->n; ++i) Df[i] = (float)Dg[i];

  const int ngc = ncomp > 1 ? dim * dim : (dim == 2 ? 3 : 6);
  const size_t gsz = (size_t)n_elem * nd * ngc * (p->geom == kG64 ? 8 : 4);
  double *d_xyz = nullptr, *d_Dg = nullptr, *d_w = nullptr, *d_Mloc = nullptr;
  int *d_bad = nullptr;
  hipStream_t st = p->stream;
#define CK(x)                                                                \
  do {                                                                       \
    if ((x) != hipSuccess) {                                                 \
      dfree(d_xyz); dfree(d_Dg); dfree(d_w); dfree(d_Mloc); dfree(d_bad);    \
      return fail(MJW_ERR_HIP, std::string("setup HIP failure at line ") +   \
                                   std::to_string(__LINE__));                \
    }                                                                        \
  } while (0)
  CK(dalloc(p, &p->d_map, hmap.size() * 4));
  CK(hipMemcpyAsync(p->d_map, hmap.data(), hmap.size() * 4, hipMemcpyHostToDevice, st));
  CK(dalloc(p, &p->d_rho, n_elem * 8));
  CK(dalloc(p, &p->d_mu, n_elem * 8));
  CK(hipMemcpyAsync(p->d_rho, rs.data(), n_elem * 8, hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(p->d_mu, ms.data(), n_elem * 8, hipMemcpyHostToDevice, st));
  if (ncomp > 1) {
    CK(dalloc(p, &p->d_lam, n_elem * 8));
    CK(hipMemcpyAsync(p->d_lam, ls.data(), n_elem * 8, hipMemcpyHostToDevice, st));
  }
  CK(dalloc(p, &p->d_D, p->n * p->n * ssize(p)));
  if (p->state == kF64)
    CK(hipMemcpyAsync(p->d_D, Dg, p->n * p->n * 8, hipMemcpyHostToDevice, st));
  else
    CK(hipMemcpyAsync(p->d_D, Df.data(), p->n * p->n * 4, hipMemcpyHostToDevice, st));
  // straight-sided hexahedra: per-element trilinear coefficients (the
  // default); the stored per-node factors are then built only when the
  // stored mode is asked for (build_stored_geometry)
  const bool tri =
      dim == 3 && trilinear_coefficients(order, n_elem, elem_nodes, node_xyz, p->perm.data(),
                                         ms.data(), 1e-11, p->hcoef);
  if (!tri) {
    p->hcoef.clear();
    CK(dalloc(p, &p->d_G, gsz));
  }
  CK(hipMalloc(&d_xyz, n_nodes * dim * 8));
  CK(hipMalloc(&d_Dg, p->n * p->n * 8));
  CK(hipMalloc(&d_w, p->n * 8));
  CK(hipMalloc(&d_Mloc, (size_t)n_elem * nd * 8));
  CK(hipMalloc(&d_bad, 4));
  CK(hipMemsetAsync(d_bad, 0, 4, st));
  std::vector<double> xyz_s((size_t)n_nodes * dim);
  for (int64_t g = 0; g < n_nodes; ++g)
    for (int a = 0; a < dim; ++a) xyz_s[(size_t)p->nperm[g] * dim + a] = node_xyz[g * dim + a];
  CK(hipMemcpyAsync(d_xyz, xyz_s.data(), n_nodes * dim * 8, hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(d_Dg, Dg, p->n * p->n * 8, hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(d_w, wg, p->n * 8, hipMemcpyHostToDevice, st));
  if (launch_geometry_full(dim, order, ncomp, n_elem, p->d_map, d_xyz, d_Dg, d_w, p->d_rho,
                           p->d_mu, nullptr, nullptr, p->geom == kG64 ? (double *)p->d_G : nullptr,
                           p->geom == kG32 ? (float *)p->d_G : nullptr, d_Mloc, d_bad, st))
    CK(hipErrorLaunchFailure);
  CK(dalloc(p, &p->d_M, n_nodes * 8));
  for (int c = 0; c < nc; ++c)
    if (launch_mass_assemble(p->d_map, d_Mloc, n_elem, dim, order, p->coff[c], p->coff[c + 1],
                             p->d_M, st))
      CK(hipErrorLaunchFailure);
  int bad = 0;
  p->mass.resize(n_nodes);
  CK(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(p->mass.data(), p->d_M, n_nodes * 8, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  dfree(d_xyz); dfree(d_Dg); dfree(d_w); dfree(d_Mloc); dfree(d_bad);
  d_xyz = d_Dg = d_w = d_Mloc = nullptr;
  d_bad = nullptr;
  if (bad) return fail(MJW_ERR_GEOMETRY, "non-positive Jacobian: degenerate or inverted element");
  if (tri) {
    const size_t gb = p->geom == kG64 ? 8 : 4;
    CK(dalloc(p, &p->d_coef, p->hcoef.size() * gb));
    CK(dalloc(p, &p->d_glxw, 2 * p->n * gb));
    CK(upload_trilinear(p, xg, wg));
    p->tri_ok = true;
    p->geo = kGeoTrilinear;
  }
  CK(dalloc(p, &p->d_Minv, n_nodes * 8));
  CK(dalloc(p, &p->d_isqM, n_nodes * 8));
  CK(finish_mass(p));
  std::vector<uint8_t> dirv(n_nodes);
  for (int64_t g = 0; g < n_nodes; ++g) dirv[p->nperm[g]] = p->dir[g];
  CK(dalloc(p, &p->d_dir, n_nodes));
  CK(hipMemcpyAsync(p->d_dir, dirv.data(), n_nodes, hipMemcpyHostToDevice, st));
  CK(dalloc(p, &p->d_nperm, n_nodes * 4));
  CK(dalloc(p, &p->d_ninv, n_nodes * 4));
  CK(hipMemcpyAsync(p->d_nperm, p->nperm.data(), n_nodes * 4, hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(p->d_ninv, p->ninv.data(), n_nodes * 4, hipMemcpyHostToDevice, st));
  // xyz_s, msv, ... are read by the copies above before they go out of scope
  CK(hipStreamSynchronize(st));
  // state
  const size_t sb = (size_t)nstate(p) * ssize(p);
  CK(dalloc(p, &p->d_U[0], sb));
  CK(dalloc(p, &p->d_U[1], sb));
  CK(dalloc(p, &p->d_tmp, sb));
  CK(dalloc(p, &p->d_tmp2, sb));
  CK(hipMemsetAsync(p->d_U[0], 0, sb, st));
  CK(hipMemsetAsync(p->d_U[1], 0, sb, st));
  CK(dalloc(p, &p->d_sdiag, n_nodes * ssize(p)));
  CK(hipMemsetAsync(p->d_sdiag, 0, n_nodes * ssize(p), st));
  CK(dalloc(p, &p->d_part, (size_t)reduction_scratch_doubles(nshots) * 8));
  CK(dalloc(p, &p->d_red, (size_t)4 * nshots * 8));
  CK(dalloc(p, &p->d_ctl, 2 * 8));
  CK(hipMemsetAsync(p->d_ctl, 0, 16, st));
  CK(hipStreamSynchronize(st));
#undef CK
  p->epb = default_tile(dim, order, ncomp, p->state);
  *out = p;
  return MJW_OK;
}

extern "C" int mjw_info(const mjw_problem *p, int64_t *info) {
  if (!p || !info) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_info: NULL");
  info[0] = p->dim; info[1] = p->order; info[2] = p->ncomp; info[3] = p->N; info[4] = p->E;
  info[5] = p->ncolor; info[6] = p->nshots; info[7] = p->prec; info[8] = (int64_t)p->bytes;
  return MJW_OK;
}

// launch shapes of the time loop's stiffness kernel: elements per block
// (single shot) or shot lanes per column (interleaved shots)
static std::vector<int> problem_tiles(const mjw_problem *p) {
  if (p->ilv)
    return p->ncomp == 3 ? tile_choices_shots_elastic(p->order, p->state, p->geom)
                         : tile_choices_shots(p->order, p->state, p->geom);
  return tile_choices(p->dim, p->order, p->ncomp);
}

extern "C" int mjw_set_tile(mjw_problem *p, int epb) {
  if (!p) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL problem");
  int &slot = p->ilv ? p->st : p->epb;
  int want = epb;
  if (epb == 0) {
    want = p->ilv ? 0 : default_tile(p->dim, p->order, p->ncomp, p->state);
  } else {
    std::vector<int> v = problem_tiles(p);
    if (std::find(v.begin(), v.end(), epb) == v.end())
      FAIL(p, MJW_ERR_ARG, "unsupported launch shape for this problem");
  }
  if (want != slot && p->gexec) {
    (void)hipGraphExecDestroy(p->gexec);
    p->gexec = nullptr;
  }
  slot = want;
  return MJW_OK;
}

extern "C" int mjw_set_xcd_order(mjw_problem *p, int nxcd) {
  if (!p || nxcd < 1 || nxcd > 64) FAIL(p, MJW_ERR_ARG, "mjw_set_xcd_order: bad args");
  if (nxcd != p->nxcd && p->gexec) {
    (void)hipGraphExecDestroy(p->gexec);
    p->gexec = nullptr;
  }
  p->nxcd = nxcd;
  return MJW_OK;
}

// the stored per-node factors of an acoustic or elastic problem with
// per-element media, built from the node positions when a problem created in
// trilinear mode is switched to the stored mode
static hipError_t build_stored_geometry(mjw_problem *p) {
  if (p->d_G) return hipSuccess;
  const int dim = p->dim, nd = p->nd;
  const int ngc = p->ncomp > 1 ? dim * dim : (dim == 2 ? 3 : 6);
  const int64_t N = p->N, E = p->E;
  double xg[kMaxN], wg[kMaxN], Dg[kMaxN * kMaxN];
  gl_rule(p->order, xg, wg);
  gl_deriv(p->order, xg, Dg);
  std::vector<double> xyz_s((size_t)N * dim);
  for (int64_t g = 0; g < N; ++g)
    for (int a = 0; a < dim; ++a) xyz_s[(size_t)p->nperm[g] * dim + a] = p->xyz[g * dim + a];
  void *G = nullptr;
  double *d_xyz = nullptr, *d_Dg = nullptr, *d_w = nullptr;
  int *d_bad = nullptr;
  hipStream_t st = p->stream;
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t x) { if (e == hipSuccess) e = x; return e == hipSuccess; };
  ok(dalloc(p, &G, (size_t)E * nd * ngc * (p->geom == kG64 ? 8 : 4)));
  ok(hipMalloc(&d_xyz, xyz_s.size() * 8));
  ok(hipMalloc(&d_Dg, p->n * p->n * 8));
  ok(hipMalloc(&d_w, p->n * 8));
  ok(hipMalloc(&d_bad, 4));
  ok(hipMemsetAsync(d_bad, 0, 4, st));
  ok(hipMemcpyAsync(d_xyz, xyz_s.data(), xyz_s.size() * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_Dg, Dg, p->n * p->n * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_w, wg, p->n * 8, hipMemcpyHostToDevice, st));
  if (e == hipSuccess &&
      launch_geometry_full(dim, p->order, p->ncomp, E, p->d_map, d_xyz, d_Dg, d_w, p->d_rho,
                           p->d_mu, nullptr, nullptr, p->geom == kG64 ? (double *)G : nullptr,
                           p->geom == kG32 ? (float *)G : nullptr, nullptr, d_bad, st))
    e = hipErrorLaunchFailure;
  ok(hipStreamSynchronize(st));
  dfree(d_xyz); dfree(d_Dg); dfree(d_w); dfree(d_bad);
  if (e != hipSuccess) {
    dfree(G);
    return e;
  }
  p->d_G = G;
  return hipSuccess;
}

extern "C" int mjw_set_geometry_mode(mjw_problem *p, int mode) {
  if (!p) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL problem");
  if (mode != MJW_GEOM_STORED && mode != MJW_GEOM_TRILINEAR)
    FAIL(p, MJW_ERR_ARG, "mjw_set_geometry_mode: bad mode");
  if (mode == MJW_GEOM_TRILINEAR && !p->tri_ok)
    FAIL(p, MJW_ERR_UNSUPPORTED,
         "trilinear geometry needs a 3-D mesh of straight-sided hexahedra");
  const int g = mode == MJW_GEOM_TRILINEAR ? kGeoTrilinear : kGeoStored;
  if (g == kGeoStored && !p->d_G) {
    HIPCHK(p, hipSetDevice(p->device));
    HIPCHK(p, build_stored_geometry(p));
  }
  if (g != p->geo) {
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    if (p->bexec) { (void)hipGraphExecDestroy(p->bexec); p->bexec = nullptr; }
  }
  p->geo = g;
  return MJW_OK;
}

extern "C" int mjw_get_geometry_mode(const mjw_problem *p, int *mode, int *trilinear_available) {
  if (!p || !mode) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL");
  *mode = p->geo == kGeoTrilinear ? MJW_GEOM_TRILINEAR : MJW_GEOM_STORED;
  if (trilinear_available) *trilinear_available = p->tri_ok ? 1 : 0;
  return MJW_OK;
}

extern "C" int mjw_get_tile(const mjw_problem *p, int *epb) {
  if (!p || !epb) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL");
  *epb = p->ilv ? p->st : p->epb;
  return MJW_OK;
}

// --------------------------------------------------------------- operators
static StiffArgs stiff_args(mjw_problem *p, int mode, const void *U, void *Y, int nshots) {
  StiffArgs a{};
  a.dim = p->dim; a.order = p->order; a.ncomp = p->ncomp; a.epb = p->epb;
  a.state = p->state; a.geom = p->geom; a.mode = mode; a.E = p->E;
  a.map = p->d_map; a.G = p->d_G; a.Dm = p->d_D; a.lam = p->d_lam; a.mu = p->d_mu;
  a.U = U; a.Y = Y; a.sdiag = p->d_sdiag; a.N = p->N; a.nshots = nshots;
  a.alpha = 1.0; a.beta = 0.0; a.nxcd = p->nxcd; a.grad = nullptr;
  a.geo = p->geo; a.coef = p->d_coef; a.glxw = p->d_glxw;
  a.ilv = p->ilv && nshots > 1;
  if (a.ilv) a.epb = p->st;
  return a;
}

// Y = beta*Y + alpha*A*U over all colours
static int apply_A(mjw_problem *p, double alpha, const void *U, double beta, void *Y, int nshots) {
  for (int c = 0; c < p->ncolor; ++c) {
    // beta is applied once per node, by its first toucher (flag in the map)
    StiffArgs a = stiff_args(p, kApply, U, Y, nshots);
    a.alpha = alpha;
    a.beta = beta;
    a.e0 = p->coff[c];
    a.e1 = p->coff[c + 1];
    int rc = launch_stiffness(a, p->stream);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int mjw_apply_kh_dev(mjw_problem *p, double alpha, const void *U, double beta,
                                void *Y) {
  if (!p || !U || !Y) FAIL(p, MJW_ERR_ARG, "mjw_apply_kh_dev: NULL");
  HIPCHK(p, hipSetDevice(p->device));
  // caller's layout and numbering -> solver's, apply, and back
  const int S = p->nshots;
  KCHK(p, launch_to_solver(p->state, S, p->ncomp, p->N, p->ilv, p->d_ninv, U, p->d_tmp, p->stream));
  if (beta != 0.0)
    KCHK(p, launch_to_solver(p->state, S, p->ncomp, p->N, p->ilv, p->d_ninv, Y, p->d_tmp2,
                             p->stream));
  else
    HIPCHK(p, hipMemsetAsync(p->d_tmp2, 0, nstate(p) * ssize(p), p->stream));
  KCHK(p, apply_A(p, -alpha, p->d_tmp, beta, p->d_tmp2, S));
  KCHK(p, launch_from_solver(p->state, S, p->ncomp, p->N, p->ilv, p->d_nperm, p->d_tmp2, Y,
                             p->stream));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  return MJW_OK;
}

extern "C" int mjw_node_order(const mjw_problem *p, int32_t *solver_index) {
  if (!p || !solver_index) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "mjw_node_order: NULL");
  std::copy(p->nperm.begin(), p->nperm.end(), solver_index);
  return MJW_OK;
}

// host FP64 vectors in the caller's layout [Sv][ncomp][N] and numbering <->
// device vectors in the solver's (state precision; Sv > 1 on a problem with
// interleaved shots uses the interleaved layout)
// FP64 device buffer of one vector ([ncomp][N], caller's numbering) through
// which host vectors cross the boundary; allocated on first use
static double *stage_buf(mjw_problem *p) {
  if (!p->d_stage &&
      dalloc(p, &p->d_stage, (size_t)p->ncomp * p->N * sizeof(double)) != hipSuccess)
    p->d_stage = nullptr;
  return p->d_stage;
}

// host vectors [Sv][ncomp][N] in the caller's numbering (FP64) <-> device
// vectors in the solver layout and precision; renumbered on the device, one
// shot at a time
static int upload_vec(mjw_problem *p, const double *h, void *dst, int Sv) {
  const int64_t per = (int64_t)p->ncomp * p->N;
  double *stage = stage_buf(p);
  if (!stage) return -1;
  const int ilv = (p->ilv && Sv > 1) ? 1 : 0;
  for (int s = 0; s < Sv; ++s) {
    int rc = (int)hipMemcpyAsync(stage, h + s * per, per * 8, hipMemcpyHostToDevice, p->stream);
    if (!rc) rc = launch_to_solver_f64(p->state, Sv, s, p->ncomp, p->N, ilv, p->d_ninv, stage, dst,
                                       p->stream);
    if (rc) return rc;
  }
  return (int)hipStreamSynchronize(p->stream);
}

static int download_vec(mjw_problem *p, const void *src, double *h, int Sv) {
  const int64_t per = (int64_t)p->ncomp * p->N;
  double *stage = stage_buf(p);
  if (!stage) return -1;
  const int ilv = (p->ilv && Sv > 1) ? 1 : 0;
  for (int s = 0; s < Sv; ++s) {
    int rc = launch_from_solver_f64(p->state, Sv, s, p->ncomp, p->N, ilv, p->d_nperm, src, stage,
                                    p->stream);
    if (!rc) rc = (int)hipMemcpyAsync(h + s * per, stage, per * 8, hipMemcpyDeviceToHost, p->stream);
    if (!rc) rc = (int)hipStreamSynchronize(p->stream);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int mjw_apply_kh(mjw_problem *p, double alpha, const double *U, double beta,
                            double *Y) {
  if (!p || !U || !Y) FAIL(p, MJW_ERR_ARG, "mjw_apply_kh: NULL");
  HIPCHK(p, hipSetDevice(p->device));
  const int64_t n = (int64_t)p->ncomp * p->N;
  KCHK(p, upload_vec(p, U, p->d_tmp, 1));
  if (beta != 0.0)
    KCHK(p, upload_vec(p, Y, p->d_tmp2, 1));
  else
    HIPCHK(p, hipMemsetAsync(p->d_tmp2, 0, n * ssize(p), p->stream));
  KCHK(p, apply_A(p, -alpha, p->d_tmp, beta, p->d_tmp2, 1));
  KCHK(p, download_vec(p, p->d_tmp2, Y, 1));
  return MJW_OK;
}

// S_h of node g in the caller's order; 0 on Dirichlet rows (S_h is assembled
// with Dirichlet coefficient 0, TimeAcoustic.cxx:966-971)
// S_h in the caller's order (0 on Dirichlet rows, S_h being assembled with
// Dirichlet coefficient 0, TimeAcoustic.cxx:966-971), or null when undamped
static const double *damp_of(const mjw_problem *p) {
  return p->damp_c.empty() ? nullptr : p->damp_c.data();
}

extern "C" int mjw_apply_dh(mjw_problem *p, double alpha, const double *U, double beta,
                            double *Y) {
  if (!p || !U || !Y) FAIL(p, MJW_ERR_ARG, "mjw_apply_dh: NULL");
  diag_axpby(p->N, p->ncomp, p->mass.data(), nullptr, 0.0, alpha, U, beta, Y);
  return MJW_OK;
}

extern "C" int mjw_apply_sh(mjw_problem *p, double alpha, const double *U, double beta,
                            double *Y) {
  if (!p || !U || !Y) FAIL(p, MJW_ERR_ARG, "mjw_apply_sh: NULL");
  if (damp_of(p))
    diag_axpby(p->N, p->ncomp, damp_of(p), nullptr, 0.0, alpha, U, beta, Y);
  else  // S = 0
    diag_axpby(p->N, p->ncomp, nullptr, nullptr, 0.0, alpha, U, beta, Y);
  return MJW_OK;
}

extern "C" int mjw_apply_dh_minus_dt_sh(mjw_problem *p, double dt, double alpha, const double *U,
                                        double beta, double *Y) {
  if (!p || !U || !Y) FAIL(p, MJW_ERR_ARG, "mjw_apply_dh_minus_dt_sh: NULL");
  if (!(dt > 0.0) || !std::isfinite(dt)) FAIL(p, MJW_ERR_ARG, "mjw_apply_dh_minus_dt_sh: bad dt");
  // coefficient mass - (dt/2) S per node
  diag_axpby(p->N, p->ncomp, p->mass.data(), damp_of(p), -0.5 * dt, alpha, U, beta, Y);
  return MJW_OK;
}

extern "C" int mjw_solve_dh(mjw_problem *p, double *Y) {
  if (!p || !Y) FAIL(p, MJW_ERR_ARG, "mjw_solve_dh: NULL");
  diag_solve(p->N, p->ncomp, p->mass.data(), nullptr, 0.0, Y);
  return MJW_OK;
}

extern "C" int mjw_solve_dh_plus_dt_sh(mjw_problem *p, double dt, double *Y) {
  if (!p || !Y) FAIL(p, MJW_ERR_ARG, "mjw_solve_dh_plus_dt_sh: NULL");
  if (!(dt > 0.0) || !std::isfinite(dt)) FAIL(p, MJW_ERR_ARG, "mjw_solve_dh_plus_dt_sh: bad dt");
  diag_solve(p->N, p->ncomp, p->mass.data(), damp_of(p), 0.5 * dt, Y);
  return MJW_OK;
}

extern "C" int mjw_set_dirichlet_condition(mjw_problem *p, double *Y) {
  if (!p || !Y) FAIL(p, MJW_ERR_ARG, "mjw_set_dirichlet_condition: NULL");
  diag_mask(p->N, p->ncomp, p->dir.data(), Y);
  return MJW_OK;
}

extern "C" int mjw_get_mass(mjw_problem *p, double *D) {
  if (!p || !D) FAIL(p, MJW_ERR_ARG, "mjw_get_mass: NULL");
  std::copy(p->mass.begin(), p->mass.end(), D);
  return MJW_OK;
}

// ------------------------------------------------------ sources, receivers
// elements whose bounding box contains x (with a relative tolerance)
static bool in_bbox(const mjw_problem *p, int64_t e, const double *x) {
  const double *b = &p->bbox[e * 2 * p->dim];
  double h = 0;
  for (int a = 0; a < p->dim; ++a) h = std::max(h, b[p->dim + a] - b[a]);
  const double tol = 1e-9 * h;
  for (int a = 0; a < p->dim; ++a)
    if (x[a] < b[a] - tol || x[a] > b[p->dim + a] + tol) return false;
  return true;
}

// deriv < 0: phi = basis values at x; deriv = a: d phi / d x_a
static bool locate_in(const mjw_problem *p, int64_t e, const double *x, double *phi,
                      int deriv = -1) {
  std::vector<double> xq((size_t)p->nd * p->dim);
  for (int q = 0; q < p->nd; ++q)
    for (int a = 0; a < p->dim; ++a)
      xq[(size_t)q * p->dim + a] = p->xyz[(int64_t)p->en[e * p->nd + q] * p->dim + a];
  double xi[3];
  if (!locate_point(p->dim, p->order, xq.data(), x, xi)) return false;
  if (deriv < 0) {
    tensor_basis(p->dim, p->order, xi, phi);
  } else {
    std::vector<double> g((size_t)p->dim * p->nd);
    tensor_grad(p->dim, p->order, xq.data(), xi, g.data());
    std::copy(g.begin() + (size_t)deriv * p->nd, g.begin() + (size_t)(deriv + 1) * p->nd, phi);
  }
  return true;
}

// buckets of about one element each per direction over the mesh box
static void build_locator(mjw_problem *p) {
  if (!p->loc_ptr.empty()) return;
  const int dim = p->dim;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t e = 0; e < p->E; ++e)
    for (int a = 0; a < dim; ++a) {
      lo[a] = std::min(lo[a], p->bbox[e * 2 * dim + a]);
      hi[a] = std::max(hi[a], p->bbox[e * 2 * dim + dim + a]);
    }
  const int per = std::max(1, (int)std::llround(std::pow((double)p->E, 1.0 / dim)));
  int64_t nb = 1;
  for (int a = 0; a < 3; ++a) {
    p->loc_n[a] = a < dim ? std::min(per, 1024) : 1;
    p->loc_lo[a] = a < dim ? lo[a] : 0.0;
    p->loc_h[a] = a < dim ? std::max((hi[a] - lo[a]) / p->loc_n[a], 1e-300) : 1.0;
    nb *= p->loc_n[a];
  }
  auto range = [&](int a, double x0, double x1, int &i0, int &i1) {
    i0 = (int)std::floor((x0 - p->loc_lo[a]) / p->loc_h[a]);
    i1 = (int)std::floor((x1 - p->loc_lo[a]) / p->loc_h[a]);
    i0 = std::max(0, std::min(p->loc_n[a] - 1, i0));
    i1 = std::max(0, std::min(p->loc_n[a] - 1, i1));
  };
  std::vector<int32_t> cnt(nb + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      for (int64_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
      p->loc_ptr = cnt;
      p->loc_elem.assign(cnt[nb], 0);
    }
    std::vector<int32_t> fill(p->loc_ptr.begin(), p->loc_ptr.end());
    for (int64_t e = 0; e < p->E; ++e) {
      const double *b = &p->bbox[e * 2 * dim];
      double h = 0;
      for (int a = 0; a < dim; ++a) h = std::max(h, b[dim + a] - b[a]);
      const double tol = 1e-9 * h;
      int r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0};
      for (int a = 0; a < dim; ++a) range(a, b[a] - tol, b[dim + a] + tol, r0[a], r1[a]);
      for (int z = r0[2]; z <= r1[2]; ++z)
        for (int y = r0[1]; y <= r1[1]; ++y)
          for (int x = r0[0]; x <= r1[0]; ++x) {
            const int64_t k = x + (int64_t)p->loc_n[0] * (y + (int64_t)p->loc_n[1] * z);
            if (pass == 0) ++cnt[k + 1];
            else p->loc_elem[fill[k]++] = (int32_t)e;
          }
    }
  }
}

// elements whose box holds x, in element order (empty outside the mesh box)
static std::pair<const int32_t *, const int32_t *> candidates(mjw_problem *p, const double *x) {
  build_locator(p);
  int64_t k = 0, stride = 1;
  for (int a = 0; a < p->dim; ++a) {
    const double t = (x[a] - p->loc_lo[a]) / p->loc_h[a];
    if (!(t >= -1e-6 && t <= p->loc_n[a] + 1e-6)) return {nullptr, nullptr};
    const int i = std::max(0, std::min(p->loc_n[a] - 1, (int)std::floor(t)));
    k += i * stride;
    stride *= p->loc_n[a];
  }
  return {p->loc_elem.data() + p->loc_ptr[k], p->loc_elem.data() + p->loc_ptr[k + 1]};
}

static int store_sources(mjw_problem *p, const std::vector<std::map<int64_t, double>> &acc);

extern "C" int mjw_set_source_vectors(mjw_problem *p, const double *F) {
  if (!p || !F) FAIL(p, MJW_ERR_ARG, "mjw_set_source_vectors: NULL");
  HIPCHK(p, hipSetDevice(p->device));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  const int64_t n1 = (int64_t)p->ncomp * p->N;
  std::vector<std::map<int64_t, double>> acc(p->nshots);
  for (int s = 0; s < p->nshots; ++s)
    for (int64_t i = 0; i < n1; ++i) {
      const double v = F[(size_t)s * n1 + i];
      if (!std::isfinite(v)) FAIL(p, MJW_ERR_ARG, "mjw_set_source_vectors: non-finite entry");
      if (v != 0.0) acc[s][i] = v;
    }
  return store_sources(p, acc);
}

extern "C" int mjw_set_sources(mjw_problem *p, int nsrc, const double *pts, const double *pol,
                               const int32_t *shot) {
  if (!p || nsrc < 0 || (nsrc > 0 && !pts)) FAIL(p, MJW_ERR_ARG, "mjw_set_sources: bad args");
  HIPCHK(p, hipSetDevice(p->device));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  // one source vector per shot, the sum over its points in the order given
  // (ComputeRightHandSide, VarInstationary.cxx:620-675, forms one vector F
  // for all points); each entry of the device list is then a distinct
  // (shot, dof), so the injection has no write conflicts
  std::vector<std::map<int64_t, double>> acc(p->nshots);
  std::vector<double> phi(p->nd);
  for (int k = 0; k < nsrc; ++k) {
    const int s = shot ? shot[k] : 0;
    if (s < 0 || s >= p->nshots) FAIL(p, MJW_ERR_ARG, "source shot index out of range");
    const double *x = pts + (size_t)k * p->dim;
    std::map<int64_t, double> b;  // node -> sum of phi over the elements holding x
    int cnt = 0;
    const auto cand = candidates(p, x);
    for (const int32_t *ce = cand.first; ce != cand.second; ++ce) {
      const int64_t e = *ce;
      if (!in_bbox(p, e, x)) continue;
      if (!locate_in(p, e, x, phi.data())) continue;
      for (int q = 0; q < p->nd; ++q) b[p->en[e * p->nd + q]] += phi[q];
      ++cnt;
    }
    if (cnt == 0) FAIL(p, MJW_ERR_ARG, "source point " + std::to_string(k) + " outside the mesh");
    // average over the elements holding the point (DefineSourceElliptic.cxx:2712-2725)
    for (int m = 0; m < p->ncomp; ++m) {
      const double pm = pol ? pol[(size_t)k * p->ncomp + m] : 1.0;
      for (auto &kv : b) {
        double v = pm * kv.second;
        if (cnt > 1) v *= 1.0 / cnt;
        acc[s][(int64_t)m * p->N + kv.first] += v;
      }
    }
  }
  return store_sources(p, acc);
}

// The device lists of the source vectors acc[shot] (caller numbering m*N + g):
// per shot, Dirichlet rows dropped, then values below eps*||F||
// (VarInstationary.cxx:637-658)
static int store_sources(mjw_problem *p, const std::vector<std::map<int64_t, double>> &acc) {
  std::vector<int64_t> dof, hdof;
  std::vector<int32_t> sh;
  std::vector<double> val;
  for (int s = 0; s < p->nshots; ++s) {
    double nrm = 0.0;
    for (auto &kv : acc[s])
      if (!p->dir[kv.first % p->N]) nrm += kv.second * kv.second;
    const double thr = 2.220446049250313e-16 * std::sqrt(nrm);
    for (auto &kv : acc[s]) {
      const int64_t g = kv.first % p->N, m = kv.first / p->N;
      if (p->dir[g] || std::fabs(kv.second) <= thr) continue;
      dof.push_back(m * p->N + p->nperm[g]);
      hdof.push_back(kv.first);
      sh.push_back(s);
      val.push_back(kv.second);
    }
  }
  p->h_src_dof = hdof;
  p->h_src_shot = sh;
  p->h_src_val = val;
  dfree(p->d_src_dof); dfree(p->d_src_shot); dfree(p->d_src_val);
  p->d_src_dof = nullptr; p->d_src_shot = nullptr; p->d_src_val = nullptr;
  p->nsrc = (int)dof.size();
  if (p->nsrc) {
    HIPCHK(p, dalloc(p, &p->d_src_dof, dof.size() * 8));
    HIPCHK(p, dalloc(p, &p->d_src_shot, sh.size() * 4));
    HIPCHK(p, dalloc(p, &p->d_src_val, val.size() * 8));
    HIPCHK(p, hipMemcpy(p->d_src_dof, dof.data(), dof.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(p, hipMemcpy(p->d_src_shot, sh.data(), sh.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(p, hipMemcpy(p->d_src_val, val.data(), val.size() * 8, hipMemcpyHostToDevice));
  }
  // both step graphs hold the old source arrays
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->bexec) { (void)hipGraphExecDestroy(p->bexec); p->bexec = nullptr; }
  return MJW_OK;
}

extern "C" int mjw_set_temporal_source(mjw_problem *p, int kind, double f0, double param,
                                       double t_init, double eps_time) {
  if (!p) FAIL(p, MJW_ERR_ARG, "mjw_set_temporal_source: NULL");
  if (kind == MJW_TIME_NONE) {
    p->has_tsrc = false;
    return MJW_OK;
  }
  if (kind < MJW_TIME_RICKER || kind > MJW_TIME_SINUS_GAUSSIAN)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_temporal_source: unknown kind");
  if (!(f0 > 0.0) || !std::isfinite(f0) || !std::isfinite(t_init) || !(eps_time >= 0.0))
    FAIL(p, MJW_ERR_ARG, "mjw_set_temporal_source: bad f0, t_init or eps");
  const double tl = time_source_limit(kind, f0, param);
  if (std::isnan(tl)) FAIL(p, MJW_ERR_ARG, "mjw_set_temporal_source: bad parameter for this kind");
  p->tsrc.init(f0, t_init, eps_time, kind, param);
  p->tlimit = tl;
  p->has_tsrc = true;
  return MJW_OK;
}

extern "C" int mjw_set_time_source(mjw_problem *p, int kind, double f0, double t_init,
                                   double eps_time) {
  if (p && kind != MJW_TIME_NONE && kind != MJW_TIME_RICKER)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_time_source: Ricker only (mjw_set_temporal_source "
                                 "takes the other kinds and their parameter)");
  return mjw_set_temporal_source(p, kind, f0, 0.0, t_init, eps_time);
}

extern "C" int mjw_add_scalar_time_source(mjw_problem *p, int shot, double alpha, double t,
                                          int nb_deriv, double *Y) {
  if (!p || !Y) FAIL(p, MJW_ERR_ARG, "mjw_add_scalar_time_source: NULL");
  if (shot < 0 || shot >= p->nshots) FAIL(p, MJW_ERR_ARG, "shot index out of range");
  if (nb_deriv < 0 || nb_deriv >= GlTimeSource::kNDer)
    FAIL(p, MJW_ERR_ARG, "mjw_add_scalar_time_source: nb_deriv out of range");
  if (!p->has_tsrc) FAIL(p, MJW_ERR_STATE, "mjw_set_time_source has not been called");
  if (t >= p->tlimit) return MJW_OK;
  const double a = alpha * p->tsrc.eval(t, nb_deriv);
  for (size_t i = 0; i < p->h_src_dof.size(); ++i)
    if (p->h_src_shot[i] == shot) Y[p->h_src_dof[i]] += a * p->h_src_val[i];
  return MJW_OK;
}

extern "C" int mjw_set_pulse_derivatives(mjw_problem *p, int nsteps, const double *d1,
                                         const double *d2) {
  if (!p || nsteps < 0 || (nsteps > 0 && (!d1 || !d2)))
    FAIL(p, MJW_ERR_ARG, "mjw_set_pulse_derivatives: bad args");
  p->h_pulse[1].assign(d1, d1 + nsteps);
  p->h_pulse[2].assign(d2, d2 + nsteps);
  p->inited = p->inited && p->torder == 2;
  return MJW_OK;
}

extern "C" int mjw_set_pulse(mjw_problem *p, int nsteps, const double *table) {
  if (!p || nsteps < 0 || (nsteps > 0 && !table)) FAIL(p, MJW_ERR_ARG, "mjw_set_pulse: bad args");
  HIPCHK(p, hipSetDevice(p->device));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  dfree(p->d_pulse);
  p->d_pulse = nullptr;
  p->npulse = nsteps;
  if (nsteps) {
    HIPCHK(p, dalloc(p, &p->d_pulse, (size_t)nsteps * 8));
    HIPCHK(p, hipMemcpy(p->d_pulse, table, (size_t)nsteps * 8, hipMemcpyHostToDevice));
  }
  p->h_pulse[0].assign(table, table + nsteps);
  p->inited = p->inited && p->torder == 2;  // the fourth-order tables are scaled at init
  // both step graphs hold the old table
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->bexec) { (void)hipGraphExecDestroy(p->bexec); p->bexec = nullptr; }
  return MJW_OK;
}

static int set_receivers(mjw_problem *p, int nrec, const double *pts, const int32_t *comp,
                         const int32_t *deriv, int64_t *elem) {
  if (!p || nrec < 0 || (nrec > 0 && !pts)) FAIL(p, MJW_ERR_ARG, "mjw_set_receivers: bad args");
  HIPCHK(p, hipSetDevice(p->device));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  std::vector<int32_t> gid((size_t)nrec * p->nd), cm(nrec, 0);
  std::vector<double> w((size_t)nrec * p->nd);
  for (int r = 0; r < nrec; ++r) {
    const double *x = pts + (size_t)r * p->dim;
    cm[r] = comp ? comp[r] : 0;
    if (cm[r] < 0 || cm[r] >= p->ncomp) FAIL(p, MJW_ERR_ARG, "receiver component out of range");
    const int dv = deriv ? deriv[r] : -1;
    if (dv < -1 || dv >= p->dim) FAIL(p, MJW_ERR_ARG, "receiver derivative out of range");
    int64_t found = -1;
    const auto cand = candidates(p, x);
    for (const int32_t *ce = cand.first; ce != cand.second && found < 0; ++ce) {
      if (!in_bbox(p, *ce, x)) continue;
      if (locate_in(p, *ce, x, &w[(size_t)r * p->nd], dv)) found = *ce;
    }
    if (found < 0) FAIL(p, MJW_ERR_ARG, "receiver " + std::to_string(r) + " outside the mesh");
    if (elem) elem[r] = found;
    for (int q = 0; q < p->nd; ++q)
      gid[(size_t)r * p->nd + q] = p->nperm[p->en[found * p->nd + q]];
  }
  dfree(p->d_rec_gid); dfree(p->d_rec_w); dfree(p->d_rec_comp);
  p->d_rec_gid = nullptr; p->d_rec_w = nullptr; p->d_rec_comp = nullptr;
  p->nrec = nrec;
  {
    // receiver transpose R^T as CSR over the unique nodes it touches (the
    // adjoint source), entries of one node in (receiver, local node) order
    std::vector<std::pair<int64_t, std::pair<int32_t, double>>> ent;
    for (int r = 0; r < nrec; ++r)
      for (int q = 0; q < p->nd; ++q) {
        const double wv = w[(size_t)r * p->nd + q];
        if (wv != 0.0)
          ent.push_back({(int64_t)cm[r] * p->N + gid[(size_t)r * p->nd + q], {r, wv}});
      }
    std::stable_sort(ent.begin(), ent.end(),
                     [](const auto &x, const auto &y) { return x.first < y.first; });
    std::vector<int32_t> node, ptr(1, 0), rec;
    std::vector<double> wt;
    for (size_t i = 0; i < ent.size(); ++i) {
      if (i == 0 || ent[i].first != ent[i - 1].first) {
        if (i) ptr.push_back((int32_t)rec.size());
        node.push_back((int32_t)ent[i].first);
      }
      rec.push_back(ent[i].second.first);
      wt.push_back(ent[i].second.second);
    }
    if (!ent.empty()) ptr.push_back((int32_t)rec.size());
    dfree(p->d_adj_node); dfree(p->d_adj_ptr); dfree(p->d_adj_rec); dfree(p->d_adj_w);
    p->d_adj_node = nullptr; p->d_adj_ptr = nullptr; p->d_adj_rec = nullptr; p->d_adj_w = nullptr;
    p->nadj = (int)node.size();
    if (p->nadj) {
      HIPCHK(p, dalloc(p, &p->d_adj_node, node.size() * 4));
      HIPCHK(p, dalloc(p, &p->d_adj_ptr, ptr.size() * 4));
      HIPCHK(p, dalloc(p, &p->d_adj_rec, rec.size() * 4));
      HIPCHK(p, dalloc(p, &p->d_adj_w, wt.size() * 8));
      HIPCHK(p, hipMemcpy(p->d_adj_node, node.data(), node.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(p, hipMemcpy(p->d_adj_ptr, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(p, hipMemcpy(p->d_adj_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(p, hipMemcpy(p->d_adj_w, wt.data(), wt.size() * 8, hipMemcpyHostToDevice));
    }
    if (p->bexec) { (void)hipGraphExecDestroy(p->bexec); p->bexec = nullptr; }
  }
  if (nrec) {
    HIPCHK(p, dalloc(p, &p->d_rec_gid, gid.size() * 4));
    HIPCHK(p, dalloc(p, &p->d_rec_w, w.size() * 8));
    HIPCHK(p, dalloc(p, &p->d_rec_comp, cm.size() * 4));
    HIPCHK(p, hipMemcpy(p->d_rec_gid, gid.data(), gid.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(p, hipMemcpy(p->d_rec_w, w.data(), w.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(p, hipMemcpy(p->d_rec_comp, cm.data(), cm.size() * 4, hipMemcpyHostToDevice));
  }
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  return MJW_OK;
}

extern "C" int mjw_set_receivers(mjw_problem *p, int nrec, const double *pts,
                                 const int32_t *comp) {
  return set_receivers(p, nrec, pts, comp, nullptr, nullptr);
}

extern "C" int mjw_set_receivers_grad(mjw_problem *p, int nrec, const double *pts,
                                      const int32_t *comp, const int32_t *deriv, int64_t *elem) {
  return set_receivers(p, nrec, pts, comp, deriv, elem);
}

// ---------------------------------------------------------------- leapfrog
static int pml_init(mjw_problem *p, double t_begin, double dt, const double *u0,
                    const double *du0);

// Fourth-order modified equation (ModifiedEquationIterator, order 4): three
// levels U^{n-1} (d_U3), U^n (d_U[cur]), U^{n+1} (d_U[cur^1]); the start
// from u0, du0 by the fourth-order Taylor expansion of SetInitialConditionS
// (ModifiedEquation.cxx:124-170): d^k U/dt^k = D^{-1}(-A d^{k-2} U - S d^{k-1} U),
// k = 2, 3, 4, each zero on the Dirichlet nodes, no source near t_begin.
static int me4_init(mjw_problem *p, double t_begin, double dt, const double *u0,
                    const double *du0) {
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  const int64_t ns = nstate(p);
  const size_t bytes = (size_t)ns * ssize(p);
  const int S = p->nshots;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }  // dt is baked in
  if (!p->d_U3) HIPCHK(p, dalloc(p, &p->d_U3, bytes));
  for (void *&v : p->d_me)
    if (!v) HIPCHK(p, dalloc(p, &v, bytes));
  // S per node (solver order, 0 on Dirichlet nodes), or none
  dfree(p->d_Sme);
  p->d_Sme = nullptr;
  if (!p->damp_S.empty()) {
    std::vector<double> sv(p->N);
    for (int64_t g = 0; g < p->N; ++g) sv[p->nperm[g]] = p->dir[g] ? 0.0 : p->damp_S[p->nperm[g]];
    HIPCHK(p, dalloc(p, &p->d_Sme, (size_t)p->N * 8));
    HIPCHK(p, hipMemcpy(p->d_Sme, sv.data(), (size_t)p->N * 8, hipMemcpyHostToDevice));
  }
  // source tables scaled by dt^2 (f' and f'' zero where not given)
  const int64_t np = p->npulse;
  dfree(p->d_pulse_me);
  p->d_pulse_me = nullptr;
  if (np) {
    std::vector<double> tab(3 * np, 0.0);
    for (int k = 0; k < 3; ++k)
      for (int64_t n = 0; n < np && n < (int64_t)p->h_pulse[k].size(); ++n)
        tab[k * np + n] = dt * dt * p->h_pulse[k][n];
    HIPCHK(p, dalloc(p, &p->d_pulse_me, tab.size() * 8));
    HIPCHK(p, hipMemcpy(p->d_pulse_me, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
  }
  KCHK(p, launch_make_sdiag(p->state, p->N, p->d_M, p->d_dir, dt * dt, p->d_sdiag, st));
  if (!p->d_sdiag4) HIPCHK(p, dalloc(p, &p->d_sdiag4, (size_t)p->N * ssize(p)));
  KCHK(p, launch_make_sdiag(p->state, p->N, p->d_M, p->d_dir, dt * dt / 12.0, p->d_sdiag4, st));
  p->dt = dt;
  p->t_begin = t_begin;
  p->zero_ic = (u0 == nullptr && du0 == nullptr);
  p->adj_ready = false;
  p->cur = 0;
  p->me_phase = 0;
  void *U1 = p->d_U[0], *U2 = p->d_U[1], *U0 = p->d_U3;
  void *V = p->d_me[1], *W = p->d_tmp2, *X = p->d_me[2], *Z = p->d_me[0];
  if (u0) KCHK(p, upload_vec(p, u0, U1, S));
  else HIPCHK(p, hipMemsetAsync(U1, 0, bytes, st));
  KCHK(p, launch_mask(p->state, ns, S, p->ilv, p->N, p->d_dir, U1, st));
  if (du0) KCHK(p, upload_vec(p, du0, V, S));
  else HIPCHK(p, hipMemsetAsync(V, 0, bytes, st));
  auto node = [&](int op, const void *a, const void *b, const void *x, const void *u,
                  const void *w, double c0, double c1, double c2, void *y) {
    return launch_me4(p->state, op, ns, S, p->ilv, p->N, p->d_M, p->d_Sme, p->d_dir, a, b, x, u,
                      w, c0, c1, c2, y, st);
  };
  // d2 = D^{-1}(-A U - S du0), d3 = D^{-1}(-A du0 - S d2), d4 = D^{-1}(-A d2 - S d3)
  KCHK(p, apply_A(p, -1.0, U1, 0.0, W, S));
  KCHK(p, node(kMeSolve, W, V, nullptr, nullptr, nullptr, 0, 0, 0, W));
  KCHK(p, apply_A(p, -1.0, V, 0.0, X, S));
  KCHK(p, node(kMeSolve, X, W, nullptr, nullptr, nullptr, 0, 0, 0, X));
  KCHK(p, apply_A(p, -1.0, W, 0.0, Z, S));
  KCHK(p, node(kMeSolve, Z, X, nullptr, nullptr, nullptr, 0, 0, 0, Z));
  // U^{+-1} = U + dt(+-du0 + dt(d2/2 + dt(+-d3/6 + dt d4/24))); the Dirichlet
  // rows of U^{-1} keep -dt du0 (SetInitialConditionS masks only U^0, U^1)
  KCHK(p, node(kMeTaylor, V, W, X, Z, U1, dt, 1.0, 1.0, U2));
  KCHK(p, node(kMeTaylor, V, W, X, Z, U1, dt, -1.0, 0.0, U0));
  p->step = 0;
  int64_t ctl[2] = {0, 0};
  HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, st));
  HIPCHK(p, hipStreamSynchronize(st));
  p->inited = true;
  return MJW_OK;
}

extern "C" int mjw_set_time_order(mjw_problem *p, int order) {
  if (!p) FAIL(p, MJW_ERR_ARG, "mjw_set_time_order: NULL");
  if (order != 2 && order != 4) FAIL(p, MJW_ERR_ARG, "mjw_set_time_order: order 2 or 4");
  if (order == 4 && p->pml)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_time_order: the PML layers run the first-order system");
  if (order != p->torder) {
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    p->torder = order;
    p->inited = false;
  }
  return MJW_OK;
}

extern "C" int mjw_leapfrog_init(mjw_problem *p, double t_begin, double dt, const double *u0,
                                 const double *du0) {
  if (!p || !(dt > 0.0) || !std::isfinite(dt)) FAIL(p, MJW_ERR_ARG, "mjw_leapfrog_init: bad dt");
  if (p->pml) return pml_init(p, t_begin, dt, u0, du0);
  if (p->torder == 4) return me4_init(p, t_begin, dt, u0, du0);
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  p->dt = dt;
  p->t_begin = t_begin;
  p->zero_ic = (u0 == nullptr && du0 == nullptr);
  p->adj_ready = false;
  const int64_t ns = nstate(p);
  const int64_t n1 = (int64_t)p->ncomp * p->N;
  if (!p->damp_S.empty()) {
    // the step's kernel list depends on the damped lists rebuilt here
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    // damped: s = dt^2 / (D + dt/2 S), k = dt S / (D + dt/2 S) on the damped
    // nodes, -dt/2 S / D for the first step's -S du0 term
    std::vector<double> meff(p->N), kk, c0;
    std::vector<int32_t> idx;
    for (int64_t g = 0; g < p->N; ++g) {
      const int64_t i = p->nperm[g];
      const double S = p->damp_S[i], M = p->mass[g];
      meff[i] = M + 0.5 * dt * S;
      if (S > 0.0 && !p->dir[g]) {
        idx.push_back((int32_t)i);
        kk.push_back(dt * S / meff[i]);
        c0.push_back(-0.5 * dt * S / M);
      }
    }
    dfree(p->d_Meff); dfree(p->d_damp_k); dfree(p->d_damp_idx); dfree(p->d_damp_save);
    p->d_Meff = nullptr; p->d_damp_k = nullptr; p->d_damp_idx = nullptr; p->d_damp_save = nullptr;
    p->ndamp = (int64_t)idx.size();
    HIPCHK(p, dalloc(p, &p->d_Meff, (size_t)p->N * 8));
    HIPCHK(p, hipMemcpy(p->d_Meff, meff.data(), (size_t)p->N * 8, hipMemcpyHostToDevice));
    KCHK(p, launch_make_sdiag(p->state, p->N, p->d_Meff, p->d_dir, dt * dt, p->d_sdiag, st));
    if (p->ndamp) {
      HIPCHK(p, dalloc(p, &p->d_damp_idx, idx.size() * 4));
      HIPCHK(p, dalloc(p, &p->d_damp_k, 2 * kk.size() * 8));
      HIPCHK(p, dalloc(p, &p->d_damp_save,
                       (size_t)p->ndamp * p->ncomp * p->nshots * ssize(p)));
      kk.insert(kk.end(), c0.begin(), c0.end());
      HIPCHK(p, hipMemcpy(p->d_damp_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
      HIPCHK(p, hipMemcpy(p->d_damp_k, kk.data(), kk.size() * 8, hipMemcpyHostToDevice));
    }
  } else {
    KCHK(p, launch_make_sdiag(p->state, p->N, p->d_M, p->d_dir, dt * dt, p->d_sdiag, st));
  }
  p->cur = 0;
  void *U0 = p->d_U[0], *U1 = p->d_U[1];
  if (u0) {
    KCHK(p, upload_vec(p, u0, U0, p->nshots));
  } else {
    HIPCHK(p, hipMemsetAsync(U0, 0, ns * ssize(p), st));
  }
  KCHK(p, launch_mask(p->state, ns, p->nshots, p->ilv, p->N, p->d_dir, U0, st));
  // U1 = ((0.5 dt) D^{-1}(-A U0) + du0) dt + U0   (ModifiedEquation.cxx:112-118)
  KCHK(p, apply_A(p, -1.0, U0, 0.0, p->d_tmp, p->nshots));
  KCHK(p, launch_scale_by(p->state, ns, p->nshots, p->ilv, p->N, p->d_Minv, 0.5 * dt, p->d_tmp,
                          st));
  if (du0) {
    KCHK(p, upload_vec(p, du0, p->d_tmp2, p->nshots));
    KCHK(p, launch_axpby(p->state, ns, 1.0, p->d_tmp2, 1.0, p->d_tmp, st));
    // damping: D U'' = -A U - S U' at t_begin (ModifiedEquation.cxx:111-113)
    if (p->ndamp)
      KCHK(p, launch_damp(p->state, 2, p->ndamp, p->ncomp, p->nshots, p->ilv, p->N,
                          p->d_damp_idx, p->d_damp_k + p->ndamp, p->d_tmp, nullptr, p->d_tmp2,
                          st));
  }
  HIPCHK(p, hipMemcpyAsync(U1, U0, ns * ssize(p), hipMemcpyDeviceToDevice, st));
  KCHK(p, launch_axpby(p->state, ns, dt, p->d_tmp, 1.0, U1, st));
  KCHK(p, launch_mask(p->state, ns, p->nshots, p->ilv, p->N, p->d_dir, U1, st));
  p->step = 0;
  int64_t ctl[2] = {0, 0};
  HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, st));
  HIPCHK(p, hipStreamSynchronize(st));
  p->inited = true;
  return MJW_OK;
}

// One fused leapfrog step n (Advance, ModifiedEquation.cxx:295-345): with
// U^n in d_U[parity] and U^{n+1} in d_U[parity^1], the colour passes apply
// the stiffness to U^{n+1} and overwrite U^n with U^{n+2}; the tail adds the
// source and records the traces of U^{n+1} (row n+1).
static int enqueue_step_pml(mjw_problem *p, double *traces);

// One step of the fourth-order modified equation (ModifiedEquationIterator::
// Advance, order 4, ModifiedEquation.cxx:347-407) with Kh = -A:
//   P = dt^2 (f(t+dt) b - A U^{n+1}),  V = dt/2 (3 U^{n+1} - 4 U^n + U^{n-1}),
//   W = D^{-1}(P - S V),  X = D^{-1}(-A V + dt^2 f'(t+dt) b - S W),
//   Z = -A W + dt^2 f''(t+dt) b - S X,
//   (D + dt/2 S) U^{n+2} = P + dt^2/12 Z + dt^2/6 S X + 2 D U^{n+1} - (D - dt/2 S) U^n,
// W, X and U^{n+2} zero on the Dirichlet nodes.  Undamped (S = 0) the V and
// X terms vanish and are skipped.  U^{n+2} overwrites U^{n-1}; the caller
// rotates the levels.  The first tail records U^{n+1} (row n+1), the last
// one advances the step counter.
static int enqueue_step_me4(mjw_problem *p, double *traces) {
  hipStream_t st = p->stream;
  const int S = p->nshots;
  const int64_t ns = nstate(p);
  const double dt = p->dt, dt2 = dt * dt;
  void *Uprev = p->d_U[p->cur], *Ucur = p->d_U[p->cur ^ 1], *Unew = p->d_U3;
  void *P = p->d_tmp, *W = p->d_tmp2, *Z = p->d_me[0], *V = p->d_me[1], *X = p->d_me[2];
  const bool damped = p->d_Sme != nullptr;
  auto node = [&](int op, const void *a, const void *b, const void *x, const void *u,
                  const void *w, double c0, double c1, double c2, void *y) {
    return launch_me4(p->state, op, ns, S, p->ilv, p->N, p->d_M, p->d_Sme, p->d_dir, a, b, x, u,
                      w, c0, c1, c2, y, st);
  };
  auto tail = [&](int k, void *into, bool record, int advance, const void *sd = nullptr) {
    TailArgs t{};
    t.state = p->state; t.ctl = p->d_ctl;
    t.pulse = p->d_pulse_me ? p->d_pulse_me + (int64_t)k * p->npulse : nullptr;
    t.npulse = p->npulse; t.nsrc = p->nsrc; t.src_dof = p->d_src_dof; t.src_shot = p->d_src_shot;
    t.src_val = p->d_src_val; t.sdiag = sd; t.nrec = record && traces ? p->nrec : 0;
    t.nd = p->nd; t.rec_gid = p->d_rec_gid; t.rec_w = p->d_rec_w; t.rec_comp = p->d_rec_comp;
    t.U = Ucur; t.P = into; t.traces = (record && traces && p->nrec) ? traces : nullptr;
    t.N = p->N; t.nshots = S; t.ilv = p->ilv; t.advance = advance; t.row_offset = 1;
    return launch_tail_n(t, p->ncomp, st);
  };
  int rc = apply_A(p, -dt2, Ucur, 0.0, P, S);
  if (!rc) rc = tail(0, P, true, 0);
  if (damped) {
    if (!rc) rc = node(kMeVel, Ucur, Uprev, Unew, nullptr, nullptr, 0.5 * dt, 0, 0, V);
    if (!rc) rc = node(kMeSolve, P, V, nullptr, nullptr, nullptr, 0, 0, 0, W);
    if (!rc) rc = apply_A(p, -1.0, V, 0.0, X, S);
    if (!rc) rc = tail(1, X, false, 0);
    if (!rc) rc = node(kMeSolve, X, W, nullptr, nullptr, nullptr, 0, 0, 0, X);
  } else {
    // undamped: U^{n+2} = W + 2 U^{n+1} - U^n - dt^2/12 D^{-1} A W
    // + dt^4/12 D^{-1} f'' b.  U^{n-1}'s storage takes W - 2 U^{n+1} + U^n;
    // the fused leapfrog pass on W with s = dt^2/(12 D) (first toucher:
    // 2 W - y - s A W, the others y - s A W) leaves the first three terms,
    // and the tail adds the source with the same s
    if (!rc) rc = node(kMeSolve, P, nullptr, nullptr, nullptr, nullptr, 0, 0, 0, W);
    if (!rc) rc = node(kMePre, W, Ucur, Uprev, nullptr, nullptr, 0, 0, 0, Unew);
    for (int c = 0; c < p->ncolor && !rc; ++c) {
      StiffArgs a = stiff_args(p, kLeap, W, Unew, S);
      a.sdiag = p->d_sdiag4;
      a.e0 = p->coff[c];
      a.e1 = p->coff[c + 1];
      rc = launch_stiffness(a, st);
    }
    if (!rc) rc = tail(2, Unew, false, 1, p->d_sdiag4);
    return rc;
  }
  if (!rc) rc = apply_A(p, -1.0, W, 0.0, Z, S);
  if (!rc) rc = tail(2, Z, false, 1);
  if (damped && !rc) rc = node(kMeSub, Z, X, nullptr, nullptr, nullptr, 0, 0, 0, Z);
  if (!rc)
    rc = node(kMeNew, P, Z, damped ? X : nullptr, Ucur, Uprev, 0.5 * dt, dt2 / 12.0, dt2 / 6.0,
              Unew);
  return rc;
}

// the levels after a fourth-order step: U^{n+1} stays, U^{n+2} (written over
// U^{n-1}) becomes d_U[cur ^ 1] and U^n becomes the oldest level
static void me4_rotate(mjw_problem *p) {
  std::swap(p->d_U[p->cur], p->d_U3);
  p->me_phase = (p->me_phase + 1) % 6;
}

static int enqueue_step(mjw_problem *p, int parity, double *traces) {
  if (p->pml) return enqueue_step_pml(p, traces);
  if (p->torder == 4) return enqueue_step_me4(p, traces);
  void *Ucur = p->d_U[parity ^ 1], *Y = p->d_U[parity];
  if (p->ndamp) {
    // keep U^n on the damped nodes before the passes overwrite it
    int rc = launch_damp(p->state, 0, p->ndamp, p->ncomp, p->nshots, p->ilv, p->N,
                         p->d_damp_idx, p->d_damp_k, Y, p->d_damp_save, nullptr, p->stream);
    if (rc) return rc;
  }
  for (int c = 0; c < p->ncolor; ++c) {
    StiffArgs a = stiff_args(p, kLeap, Ucur, Y, p->nshots);
    a.e0 = p->coff[c];
    a.e1 = p->coff[c + 1];
    int rc = launch_stiffness(a, p->stream);
    if (rc) return rc;
  }
  TailArgs t{};
  t.state = p->state; t.ctl = p->d_ctl; t.pulse = p->d_pulse; t.npulse = p->npulse;
  t.nsrc = p->nsrc; t.src_dof = p->d_src_dof; t.src_shot = p->d_src_shot; t.src_val = p->d_src_val;
  t.sdiag = p->d_sdiag; t.nrec = traces ? p->nrec : 0; t.nd = p->nd; t.rec_gid = p->d_rec_gid;
  t.rec_w = p->d_rec_w; t.rec_comp = p->d_rec_comp; t.U = Ucur; t.P = Y;
  t.traces = (traces && p->nrec) ? traces : nullptr; t.N = p->N; t.nshots = p->nshots; t.ilv = p->ilv;
  t.advance = 1;
  t.row_offset = 1;
  int rc = launch_tail_n(t, p->ncomp, p->stream);
  if (rc || !p->ndamp) return rc;
  return launch_damp(p->state, 1, p->ndamp, p->ncomp, p->nshots, p->ilv, p->N, p->d_damp_idx,
                     p->d_damp_k, Y, p->d_damp_save, Ucur, p->stream);
}

// traces of U^n (d_U[cur]) into row step - base (no source, no step change)
static int enqueue_record(mjw_problem *p, double *traces) {
  TailArgs t{};
  t.state = p->state; t.ctl = p->d_ctl; t.nrec = p->nrec; t.nd = p->nd; t.rec_gid = p->d_rec_gid;
  t.rec_w = p->d_rec_w; t.rec_comp = p->d_rec_comp; t.U = p->d_U[p->cur]; t.P = nullptr;
  t.traces = traces; t.N = p->N; t.nshots = p->nshots; t.ilv = p->ilv; t.advance = 0; t.row_offset = 0;
  return launch_tail_n(t, p->ncomp, p->stream);
}

static int ensure_traces(mjw_problem *p, int64_t nsteps) {
  const int64_t need = (nsteps + 1) * p->nshots * std::max(p->nrec, 1);
  if (need <= p->traces_cap) return 0;
  dfree(p->d_traces);
  p->d_traces = nullptr;
  p->traces_cap = 0;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (hipMalloc(&p->d_traces, need * 8) != hipSuccess) return -1;
  p->traces_cap = need;
  return 0;
}

static const int kGraphSteps = 16;  // even
// fourth-order scheme: a multiple of the 6-step cycle of its three levels
static const int kGraphSteps4 = 24;

static int graph_steps(const mjw_problem *p) { return p->torder == 4 ? kGraphSteps4 : kGraphSteps; }

static int build_graph(mjw_problem *p, bool with_traces) {
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return (int)e;
  int rc = 0;
  const int gs = graph_steps(p);
  int s = 0;
  for (; s < gs && !rc; ++s) {
    if (p->torder == 4) {
      // the captured steps rotate the levels as the eager ones do; after
      // the whole cycle the pointers are back where they started
      rc = enqueue_step(p, p->cur, with_traces ? p->d_traces : nullptr);
      me4_rotate(p);
      p->cur ^= 1;
    } else {
      rc = enqueue_step(p, (p->cur + s) & 1, with_traces ? p->d_traces : nullptr);
    }
  }
  // a capture stopped part-way has run nothing: finish the cycle of the
  // levels so that they are where they were
  for (; p->torder == 4 && s < gs; ++s) {
    me4_rotate(p);
    p->cur ^= 1;
  }
  e = hipStreamEndCapture(p->stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return (int)e;
  e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { p->gexec = nullptr; return (int)e; }
  p->g_steps = gs;
  p->g_parity = p->cur;
  p->g_traces = p->d_traces;
  p->g_with_traces = with_traces;
  return 0;
}

extern "C" int mjw_leapfrog_advance(mjw_problem *p, int nsteps, double *traces, int use_graph) {
  if (!p || nsteps < 0) FAIL(p, MJW_ERR_ARG, "mjw_leapfrog_advance: bad args");
  if (!p->inited) FAIL(p, MJW_ERR_STATE, "mjw_leapfrog_init has not been called");
  if (nsteps == 0) return MJW_OK;
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  const bool want = traces && p->nrec > 0;
  if (want) KCHK(p, ensure_traces(p, nsteps));
  int64_t ctl[2] = {p->step, p->step};
  HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, st));
  hipEvent_t ev0, ev1;
  HIPCHK(p, hipEventCreate(&ev0));
  HIPCHK(p, hipEventCreate(&ev1));
  // row 0 = U^n (the first-order system records U^n in each step's tail)
  if (want && !p->pml) KCHK(p, enqueue_record(p, p->d_traces));
  HIPCHK(p, hipEventRecord(ev0, st));
  int done = 0;
  const int gs = graph_steps(p);
  auto eager = [&]() -> int {
    int rc = enqueue_step(p, p->cur, want ? p->d_traces : nullptr);
    if (p->torder == 4) me4_rotate(p);
    p->cur ^= 1;
    return rc;
  };
  if (use_graph && nsteps >= gs) {
    // graphs are captured at even parity (fourth order: at the start of the
    // cycle of its levels)
    while (nsteps - done > 0 && (p->cur != 0 || (p->torder == 4 && p->me_phase != 0))) {
      KCHK(p, eager());
      ++done;
    }
    if (nsteps - done >= gs) {
      if (!p->gexec || p->g_parity != p->cur || p->g_with_traces != want ||
          (want && p->g_traces != p->d_traces))
        KCHK(p, build_graph(p, want));
      while (nsteps - done >= gs) {
        HIPCHK(p, hipGraphLaunch(p->gexec, st));
        done += gs;
      }
    }
  }
  for (; done < nsteps; ++done) KCHK(p, eager());
  if (p->pml) p->cur = 0;  // one state buffer
  HIPCHK(p, hipEventRecord(ev1, st));
  HIPCHK(p, hipEventSynchronize(ev1));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  p->ms_step = ms / nsteps;
  p->step += nsteps;
  if (want)
    HIPCHK(p, hipMemcpy(traces, p->d_traces, (size_t)nsteps * p->nshots * p->nrec * 8,
                        hipMemcpyDeviceToHost));
  return MJW_OK;
}

extern "C" int mjw_leapfrog_advance_static(mjw_problem *p, int nsteps, double *traces) {
  if (!p || nsteps < 0) FAIL(p, MJW_ERR_ARG, "mjw_leapfrog_advance_static: bad args");
  if (!p->inited) FAIL(p, MJW_ERR_STATE, "mjw_leapfrog_init has not been called");
  if (nsteps == 0) return MJW_OK;
  const bool want = traces && p->nrec > 0;
  const int64_t need = ((int64_t)nsteps + 1) * p->nshots * std::max(p->nrec, 1);
  if (want && need > kStaticTraceDoubles)
    FAIL(p, MJW_ERR_ARG,
         "mjw_leapfrog_advance_static: (nsteps + 1) * nshots * nrec = " + std::to_string(need) +
             " trace values exceed the static staging buffer of " +
             std::to_string(kStaticTraceDoubles) + " (advance in shorter calls)");
  HIPCHK(p, hipSetDevice(p->device));
  double *buf = nullptr;
  if (want) {
    buf = scratch_of(p->device);
    if (!buf) FAIL(p, MJW_ERR_STATE, "static trace buffer not set up on this device");
  }
  hipStream_t st = p->stream;
  KCHK(p, launch_set_ctl(p->d_ctl, p->step, p->step, st));
  if (want && !p->pml) KCHK(p, enqueue_record(p, buf));  // row 0 = U^n
  for (int k = 0; k < nsteps; ++k) {
    KCHK(p, enqueue_step(p, p->cur, buf));
    if (p->torder == 4) me4_rotate(p);
    p->cur ^= 1;
  }
  if (p->pml) p->cur = 0;
  p->step += nsteps;
  if (want)
    HIPCHK(p, hipMemcpyAsync(traces, buf, (size_t)nsteps * p->nshots * p->nrec * 8,
                             hipMemcpyDeviceToHost, st));
  HIPCHK(p, hipStreamSynchronize(st));
  return MJW_OK;
}

extern "C" int mjw_record(mjw_problem *p, double *traces) {
  if (!p || !traces) FAIL(p, MJW_ERR_ARG, "mjw_record: NULL");
  if (!p->inited) FAIL(p, MJW_ERR_STATE, "not initialised");
  if (p->nrec == 0) return MJW_OK;
  HIPCHK(p, hipSetDevice(p->device));
  KCHK(p, ensure_traces(p, 1));
  int64_t ctl[2] = {p->step, p->step};
  HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, p->stream));
  KCHK(p, enqueue_record(p, p->d_traces));
  HIPCHK(p, hipMemcpyAsync(traces, p->d_traces, (size_t)p->nshots * p->nrec * 8,
                           hipMemcpyDeviceToHost, p->stream));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  return MJW_OK;
}

static int set_probes(mjw_problem *p, int npts, const double *pts, const int32_t *deriv,
                      uint8_t *inside, int64_t *elem) {
  if (!p || npts < 0 || (npts > 0 && !pts)) FAIL(p, MJW_ERR_ARG, "mjw_set_probes: bad args");
  HIPCHK(p, hipSetDevice(p->device));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  std::vector<int32_t> gid((size_t)npts * p->nd, 0);
  std::vector<double> w((size_t)npts * p->nd, 0.0);
  for (int r = 0; r < npts; ++r) {
    const double *x = pts + (size_t)r * p->dim;
    const int dv = deriv ? deriv[r] : -1;
    if (dv < -1 || dv >= p->dim) FAIL(p, MJW_ERR_ARG, "probe derivative out of range");
    int64_t found = -1;
    const auto cand = candidates(p, x);
    for (const int32_t *ce = cand.first; ce != cand.second && found < 0; ++ce) {
      if (!in_bbox(p, *ce, x)) continue;
      if (locate_in(p, *ce, x, &w[(size_t)r * p->nd], dv)) found = *ce;
    }
    if (elem) elem[r] = found;
    if (found >= 0) {
      for (int q = 0; q < p->nd; ++q)
        gid[(size_t)r * p->nd + q] = p->nperm[p->en[found * p->nd + q]];
    } else {
      std::fill(w.begin() + (size_t)r * p->nd, w.begin() + (size_t)(r + 1) * p->nd, 0.0);
    }
    if (inside) inside[r] = found >= 0 ? 1 : 0;
  }
  dfree(p->d_pr_gid); dfree(p->d_pr_w); dfree(p->d_pr_out);
  p->d_pr_gid = nullptr; p->d_pr_w = nullptr; p->d_pr_out = nullptr;
  p->nprobe = npts;
  if (npts == 0) return MJW_OK;
  HIPCHK(p, dalloc(p, &p->d_pr_gid, gid.size() * 4));
  HIPCHK(p, dalloc(p, &p->d_pr_w, w.size() * 8));
  HIPCHK(p, dalloc(p, &p->d_pr_out, (size_t)npts * p->nshots * p->ncomp * 8));
  HIPCHK(p, hipMemcpy(p->d_pr_gid, gid.data(), gid.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(p, hipMemcpy(p->d_pr_w, w.data(), w.size() * 8, hipMemcpyHostToDevice));
  return MJW_OK;
}

extern "C" int mjw_set_probes(mjw_problem *p, int npts, const double *pts, uint8_t *inside) {
  return set_probes(p, npts, pts, nullptr, inside, nullptr);
}

extern "C" int mjw_set_probes_grad(mjw_problem *p, int npts, const double *pts,
                                   const int32_t *deriv, uint8_t *inside, int64_t *elem) {
  return set_probes(p, npts, pts, deriv, inside, elem);
}

// damping matrix S_h = sum_e sigma_e J w phi_i (lumped like the mass; the
// damping term of rho u'' + sigma u' - div(mu grad u) = f, VarHelmholtz.cxx
// :64-66,108-109 "Glob_matMass_DhSigma", assembled by TimeAcoustic.cxx:970-1004;
// elastic: sigma I in the mass tensor, VarElastic.cxx:59-71, so every
// component of a node takes the same S, assembled by TimeElastic.cxx:296-330)
// sigma per element (nodal = false) or per quadrature point [E][nd] (nodal)
static int lumped_integral(mjw_problem *p, const double *vals, bool nodal,
                           std::vector<double> &out);

static int set_damping(mjw_problem *p, const double *sigma, bool nodal) {
  if (p && p->pml) FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_damping: physical damping with PML layers is not supported");
  if (!p || !sigma) FAIL(p, MJW_ERR_ARG, "mjw_set_damping: NULL");
  HIPCHK(p, hipSetDevice(p->device));
  const int64_t E = p->E, N = p->N, per = nodal ? p->nd : 1;
  std::vector<double> ss((size_t)E * per);
  bool any = false;
  for (int64_t es = 0; es < E; ++es)
    for (int64_t q = 0; q < per; ++q) {
      const double v = sigma[p->perm[es] * per + q];
      if (!(v >= 0.0) || !std::isfinite(v))
        FAIL(p, MJW_ERR_ARG, "mjw_set_damping: sigma must be finite, >= 0");
      ss[(size_t)es * per + q] = v;
      any = any || v > 0.0;
    }
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  p->inited = false;
  p->ndamp = 0;
  if (!any) {
    p->damp_S.clear();
    p->damp_c.clear();
    return MJW_OK;
  }
  std::vector<double> S;
  const int rc = lumped_integral(p, ss.data(), nodal, S);
  if (rc) {
    p->damp_S.clear();
    p->damp_c.clear();
    return rc;
  }
  p->damp_S = S;
  p->damp_c.assign(N, 0.0);
  for (int64_t g = 0; g < N; ++g)
    if (!p->dir[g]) p->damp_c[g] = p->damp_S[p->nperm[g]];
  return MJW_OK;
}

// out[i] = sum_e sum_q v J w phi_i(x_q) (solver order): v per element
// (nodal = false) or per quadrature point [E][nd] (sorted element order,
// local order), through the geometry kernel with v in the place of rho
static int lumped_integral(mjw_problem *p, const double *vals, bool nodal,
                           std::vector<double> &out) {
  const int64_t E = p->E, N = p->N, per = nodal ? p->nd : 1;
  double xg[kMaxN], wg[kMaxN], Dg[kMaxN * kMaxN];
  gl_rule(p->order, xg, wg);
  gl_deriv(p->order, xg, Dg);
  std::vector<double> xyz_s((size_t)N * p->dim);
  for (int64_t g = 0; g < N; ++g)
    for (int a = 0; a < p->dim; ++a) xyz_s[(size_t)p->nperm[g] * p->dim + a] = p->xyz[g * p->dim + a];
  double *d_xyz = nullptr, *d_Dg = nullptr, *d_w = nullptr, *d_Mloc = nullptr, *d_sig = nullptr,
         *d_S = nullptr;
  int *d_bad = nullptr;
  hipStream_t st = p->stream;
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t x) { if (e == hipSuccess) e = x; return e == hipSuccess; };
  ok(hipMalloc(&d_xyz, xyz_s.size() * 8));
  ok(hipMalloc(&d_Dg, p->n * p->n * 8));
  ok(hipMalloc(&d_w, p->n * 8));
  ok(hipMalloc(&d_Mloc, (size_t)E * p->nd * 8));
  ok(hipMalloc(&d_sig, (size_t)E * per * 8));
  ok(hipMalloc(&d_S, (size_t)N * 8));
  ok(hipMalloc(&d_bad, 4));
  ok(hipMemsetAsync(d_bad, 0, 4, st));
  ok(hipMemcpyAsync(d_xyz, xyz_s.data(), xyz_s.size() * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_Dg, Dg, p->n * p->n * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_w, wg, p->n * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_sig, vals, (size_t)E * per * 8, hipMemcpyHostToDevice, st));
  // sigma in the place of rho: the kernel's "mass" output is sigma J w
  if (e == hipSuccess &&
      launch_geometry_full(p->dim, p->order, 1, E, p->d_map, d_xyz, d_Dg, d_w,
                           nodal ? p->d_rho : d_sig, p->d_mu, nodal ? d_sig : nullptr, nullptr,
                           nullptr, nullptr, d_Mloc, d_bad, st))
    e = hipErrorLaunchFailure;
  for (int c = 0; c < p->ncolor && e == hipSuccess; ++c)
    if (launch_mass_assemble(p->d_map, d_Mloc, E, p->dim, p->order, p->coff[c], p->coff[c + 1],
                             d_S, st))
      e = hipErrorLaunchFailure;
  out.assign(N, 0.0);
  ok(hipMemcpyAsync(out.data(), d_S, (size_t)N * 8, hipMemcpyDeviceToHost, st));
  ok(hipStreamSynchronize(st));
  dfree(d_xyz); dfree(d_Dg); dfree(d_w); dfree(d_Mloc); dfree(d_sig); dfree(d_S); dfree(d_bad);
  if (e != hipSuccess) FAIL(p, MJW_ERR_HIP, std::string("lumped integral: ") + hipGetErrorString(e));
  return MJW_OK;
}

extern "C" int mjw_lumped_integral(mjw_problem *p, const double *f_q, double *out) {
  if (!p || !f_q || !out) FAIL(p, MJW_ERR_ARG, "mjw_lumped_integral: NULL");
  HIPCHK(p, hipSetDevice(p->device));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  const int64_t E = p->E, nd = p->nd;
  // caller's element order -> sorted (colour) order
  std::vector<double> v((size_t)E * nd);
  for (int64_t es = 0; es < E; ++es)
    for (int64_t q = 0; q < nd; ++q) {
      const double x = f_q[p->perm[es] * nd + q];
      if (!std::isfinite(x)) FAIL(p, MJW_ERR_ARG, "mjw_lumped_integral: non-finite value");
      v[(size_t)es * nd + q] = x;
    }
  std::vector<double> b;
  const int rc = lumped_integral(p, v.data(), true, b);
  if (rc) return rc;
  for (int64_t g = 0; g < p->N; ++g) out[g] = b[p->nperm[g]];
  return MJW_OK;
}

extern "C" int mjw_set_damping(mjw_problem *p, const double *sigma) {
  return set_damping(p, sigma, false);
}

extern "C" int mjw_set_damping_nodal(mjw_problem *p, const double *sigma_q) {
  return set_damping(p, sigma_q, true);
}

extern "C" int mjw_probe(mjw_problem *p, double *values) {
  if (!p || !values) FAIL(p, MJW_ERR_ARG, "mjw_probe: NULL");
  if (p->nprobe == 0) return MJW_OK;
  HIPCHK(p, hipSetDevice(p->device));
  KCHK(p, launch_probe(p->state, p->d_U[p->cur], p->ilv, p->nshots, p->ncomp, p->N, p->nprobe,
                       p->nd, p->d_pr_gid, p->d_pr_w, p->d_pr_out, p->stream));
  HIPCHK(p, hipMemcpyAsync(values, p->d_pr_out, (size_t)p->nprobe * p->nshots * p->ncomp * 8,
                           hipMemcpyDeviceToHost, p->stream));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  return MJW_OK;
}

extern "C" int mjw_get_state(mjw_problem *p, double *Un, double *Un1) {
  if (!p) FAIL(p, MJW_ERR_ARG, "NULL problem");
  HIPCHK(p, hipSetDevice(p->device));
  if (Un1 && p->pml)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_get_state: the first-order system (PML layers) holds "
                                 "u^n only (pass Un1 = NULL)");
  if (Un) KCHK(p, download_vec(p, p->d_U[p->cur], Un, p->nshots));
  if (Un1) KCHK(p, download_vec(p, p->d_U[p->cur ^ 1], Un1, p->nshots));
  return MJW_OK;
}

extern "C" int mjw_state_dev(mjw_problem *p, void **Un, void **Un1) {
  if (!p) FAIL(p, MJW_ERR_ARG, "NULL problem");
  if (Un) *Un = p->d_U[p->cur];
  if (Un1) *Un1 = p->d_U[p->cur ^ 1];
  return MJW_OK;
}

extern "C" int mjw_get_step(const mjw_problem *p, int64_t *n) {
  if (!p || !n) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL");
  *n = p->step;
  return MJW_OK;
}

extern "C" int mjw_last_timing(const mjw_problem *p, double *ms2) {
  if (!p || !ms2) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL");
  ms2[0] = p->ms_step;
  ms2[1] = p->ms_stiff;
  return MJW_OK;
}

extern "C" int mjw_profile(mjw_problem *p, int nsteps, double *out) {
  if (p && p->pml) FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_profile: not available with PML layers");
  if (p && p->torder == 4)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_profile: times the second-order step only");
  if (!p || !out || nsteps < 1) FAIL(p, MJW_ERR_ARG, "mjw_profile: bad args");
  if (!p->inited) FAIL(p, MJW_ERR_STATE, "not initialised");
  if (!p->damp_S.empty()) FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_profile: undamped problems only");
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  // events only around each step's colour passes and its tail: the colour
  // passes of one step run back to back, as in the time loop
  const int nev = 2 * nsteps + 1;
  std::vector<hipEvent_t> ev(nev);
  for (auto &e : ev) HIPCHK(p, hipEventCreate(&e));
  int64_t ctl[2] = {p->step, p->step};
  HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, st));
  int k = 0;
  HIPCHK(p, hipEventRecord(ev[k++], st));
  for (int s = 0; s < nsteps; ++s) {
    void *U = p->d_U[p->cur ^ 1], *Y = p->d_U[p->cur];
    for (int c = 0; c < p->ncolor; ++c) {
      StiffArgs a = stiff_args(p, kLeap, U, Y, p->nshots);
      a.e0 = p->coff[c];
      a.e1 = p->coff[c + 1];
      KCHK(p, launch_stiffness(a, st));
    }
    HIPCHK(p, hipEventRecord(ev[k++], st));
    TailArgs t{};
    t.state = p->state; t.ctl = p->d_ctl; t.pulse = p->d_pulse; t.npulse = p->npulse;
    t.nsrc = p->nsrc; t.src_dof = p->d_src_dof; t.src_shot = p->d_src_shot;
    t.src_val = p->d_src_val; t.sdiag = p->d_sdiag; t.nrec = 0; t.nd = p->nd; t.U = U; t.P = Y;
    t.N = p->N; t.nshots = p->nshots; t.ilv = p->ilv; t.advance = 1; t.row_offset = 1;
    KCHK(p, launch_tail_n(t, p->ncomp, st));
    HIPCHK(p, hipEventRecord(ev[k++], st));
    p->cur ^= 1;
  }
  HIPCHK(p, hipEventSynchronize(ev[k - 1]));
  double stiff = 0, tail = 0;
  for (int s = 0; s < nsteps; ++s) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]);
    stiff += ms;
    (void)hipEventElapsedTime(&ms, ev[2 * s + 1], ev[2 * s + 2]);
    tail += ms;
  }
  float tot = 0;
  (void)hipEventElapsedTime(&tot, ev[0], ev[nev - 1]);
  for (auto &e : ev) (void)hipEventDestroy(e);
  p->step += nsteps;
  out[0] = stiff / (nsteps * (double)p->ncolor);
  out[1] = tail / nsteps;
  out[2] = tot / nsteps;
  out[3] = p->ncolor + 1;
  p->ms_stiff = stiff / nsteps;
  return MJW_OK;
}

extern "C" int mjw_energy(mjw_problem *p, double *energy) {
  if (p && p->pml) FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_energy: not available for the first-order system (PML layers)");
  if (p && p->torder == 4)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_energy: the discrete energy of the second-order scheme only");
  if (!p || !energy) FAIL(p, MJW_ERR_ARG, "mjw_energy: NULL");
  if (!p->inited) FAIL(p, MJW_ERR_STATE, "not initialised");
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  const int S = p->nshots;
  void *Un = p->d_U[p->cur], *Un1 = p->d_U[p->cur ^ 1];
  KCHK(p, apply_A(p, 1.0, Un, 0.0, p->d_tmp, S));
  KCHK(p, launch_dot_part(p->state, p->N, p->ncomp, S, p->ilv, nullptr, Un1, p->d_tmp, p->d_red,
                          p->d_part, st));
  KCHK(p, launch_dot_diff_part(p->state, p->N, p->ncomp, S, p->ilv, p->d_M, Un1, Un, p->d_red + S,
                               p->d_part, st));
  std::vector<double> h(2 * S);
  HIPCHK(p, hipMemcpyAsync(h.data(), p->d_red, 2 * S * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(p, hipStreamSynchronize(st));
  for (int s = 0; s < S; ++s) {
    const double e = 0.5 * h[S + s] / (p->dt * p->dt) + 0.5 * h[s];
    energy[s] = std::isfinite(e) ? e : std::numeric_limits<double>::quiet_NaN();
  }
  return MJW_OK;
}

// largest eigenvalue of a symmetric tridiagonal matrix (Sturm bisection)
static double tri_max(const std::vector<double> &al, const std::vector<double> &be) {
  const int m = (int)al.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(be[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(be[i]) : 0.0);
    lo = std::min(lo, al[i] - r);
    hi = std::max(hi, al[i] + r);
  }
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    int below = 0;
    double d = 1.0;
    for (int i = 0; i < m; ++i) {
      d = al[i] - mid - (i > 0 ? be[i - 1] * be[i - 1] / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) ++below;
    }
    if (below == m) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

extern "C" int mjw_lambda_max(mjw_problem *p, int iters, uint64_t seed, double *lambda_max) {
  if (!p || !lambda_max || iters < 2) FAIL(p, MJW_ERR_ARG, "mjw_lambda_max: bad args");
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  const int64_t n1 = (int64_t)p->ncomp * p->N;
  // Lanczos on D^{-1/2} A D^{-1/2} restricted to free nodes, single vector,
  // vectors in the state precision, scalars in FP64.
  void *q = nullptr, *qo = nullptr, *w = nullptr, *t = nullptr;
  const size_t b = n1 * ssize(p);
  HIPCHK(p, hipMalloc(&q, b));
  HIPCHK(p, hipMalloc(&qo, b));
  HIPCHK(p, hipMalloc(&w, b));
  HIPCHK(p, hipMalloc(&t, b));
  std::vector<double> h0(n1);
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + 12345;
  for (int64_t i = 0; i < n1; ++i) {
    z ^= z << 13; z ^= z >> 7; z ^= z << 17;
    h0[i] = p->dir[i % p->N] ? 0.0 : ((double)(z >> 11) / 9007199254740992.0 - 0.5);
  }
  double nrm = 0;
  for (double v : h0) nrm += v * v;
  nrm = std::sqrt(nrm);
  for (double &v : h0) v /= nrm;
  int rc = upload_vec(p, h0.data(), q, 1);
  if (!rc) rc = (int)hipMemsetAsync(qo, 0, b, st);
  std::vector<double> al, be;
  double beta = 0.0, red[2];
  for (int k = 0; k < iters && !rc; ++k) {
    rc = (int)hipMemcpyAsync(t, q, b, hipMemcpyDeviceToDevice, st);
    if (!rc) rc = launch_scale_by(p->state, n1, 1, 0, p->N, p->d_isqM, 1.0, t, st);
    if (!rc) rc = apply_A(p, 1.0, t, 0.0, w, 1);
    if (!rc) rc = launch_scale_by(p->state, n1, 1, 0, p->N, p->d_isqM, 1.0, w, st);
    if (!rc) rc = launch_dot_part(p->state, p->N, p->ncomp, 1, 0, nullptr, w, q, p->d_red, p->d_part, st);
    if (!rc) rc = (int)hipMemcpyAsync(red, p->d_red, 8, hipMemcpyDeviceToHost, st);
    if (!rc) rc = (int)hipStreamSynchronize(st);
    if (rc) break;
    const double alpha = red[0];
    al.push_back(alpha);
    // w = w - alpha q - beta qo
    rc = launch_axpby(p->state, n1, -alpha, q, 1.0, w, st);
    if (!rc) rc = launch_axpby(p->state, n1, -beta, qo, 1.0, w, st);
    if (!rc) rc = launch_dot_part(p->state, p->N, p->ncomp, 1, 0, nullptr, w, w, p->d_red, p->d_part, st);
    if (!rc) rc = (int)hipMemcpyAsync(red, p->d_red, 8, hipMemcpyDeviceToHost, st);
    if (!rc) rc = (int)hipStreamSynchronize(st);
    if (rc) break;
    beta = std::sqrt(red[0]);
    if (k + 1 == iters || beta < 1e-300) break;
    be.push_back(beta);
    std::swap(q, qo);  // qo <- q
    rc = launch_axpby(p->state, n1, 1.0 / beta, w, 0.0, q, st);
  }
  dfree(q); dfree(qo); dfree(w); dfree(t);
  if (rc) FAIL(p, MJW_ERR_HIP, "Lanczos iteration failed");
  be.resize(al.size() > 0 ? al.size() - 1 : 0);
  *lambda_max = tri_max(al, be);
  return MJW_OK;
}

// ------------------------------------------------------------ adjoint
// Discrete adjoint of the leapfrog recurrence (zero initial state):
//   U^{k+1} = 2U^k - U^{k-1} - S A U^k + S f_{k-1} b,   S = dt^2 M^{-1}
// gives, for nu_k = S lambda_k and h^k = R^T dJ/d(trace_k),
//   nu_{N-1} = S h^N, nu_N = 0,  nu_{j-1} = 2nu_j - nu_{j+1} - S A nu_j + S h^j,
//   dJ/dmu_e = - sum_{j=1}^{N-1} nu_j^T (A_e / mu_e) U^j.
// The recurrence is time reversible: U^{j-1} = 2U^j - U^{j+1} - S A U^j +
// S f_{j-1} b, so the backward pass rebuilds U^j in reverse alongside nu_j in
// the same fused two-vector kernel (geometry read once for both) and
// re-anchors U on the checkpoint pairs stored by the forward pass.

extern "C" int mjw_set_mu(mjw_problem *p, const double *mu) {
  if (p && p->pml) FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_mu: not available with PML layers");
  if (!p || !mu) FAIL(p, MJW_ERR_ARG, "mjw_set_mu: NULL");
  if (p->ncomp != 1) FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_mu: acoustic problems only");
  if (p->nodal_media)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_mu: the problem has per-quadrature-point media");
  HIPCHK(p, hipSetDevice(p->device));
  std::vector<double> old(p->E), ms(p->E), sc(p->E);
  HIPCHK(p, hipMemcpy(old.data(), p->d_mu, p->E * 8, hipMemcpyDeviceToHost));
  for (int64_t es = 0; es < p->E; ++es) {
    ms[es] = mu[p->perm[es]];
    if (!(ms[es] > 0) || !std::isfinite(ms[es])) FAIL(p, MJW_ERR_ARG, "mu must be finite, > 0");
    sc[es] = ms[es] / old[es];
  }
  double *d_sc = nullptr;
  HIPCHK(p, hipMalloc(&d_sc, p->E * 8));
  HIPCHK(p, hipMemcpy(d_sc, sc.data(), p->E * 8, hipMemcpyHostToDevice));
  const int ngc = p->dim == 2 ? 3 : 6;
  const int K = p->dim == 2 ? 1 : p->n;
  // the stored factors, when they are built, scale with mu
  int rc = p->d_G ? launch_scale_planes(p->geom == kG64 ? 0 : 1, p->d_G, (int64_t)K * ngc, p->E,
                                        p->P, d_sc, p->stream)
                  : 0;
  if (!rc) rc = (int)hipMemcpyAsync(p->d_mu, ms.data(), p->E * 8, hipMemcpyHostToDevice, p->stream);
  if (!rc) rc = (int)hipStreamSynchronize(p->stream);
  if (!rc && p->tri_ok) {
    for (int64_t es = 0; es < p->E; ++es) p->hcoef[(size_t)es * kCoef + 21] = ms[es];
    rc = (int)upload_trilinear(p, nullptr, nullptr);
  }
  dfree(d_sc);
  if (rc) FAIL(p, MJW_ERR_HIP, "mjw_set_mu: rescale failed");
  p->adj_ready = false;
  return MJW_OK;
}

// Media varying inside the elements: rho and mu at every quadrature point
// (PhysicalMedia ... SINUS / MESH / REGULAR_GRID evaluated by the caller,
// as VarGeometryProblem.cxx:1395-1500 stores them for
// VarHelmholtz.cxx:73-117).  Rebuilds the stored geometric factors and the
// lumped mass from the vertices; the per-element trilinear form no longer
// applies, so the problem switches to stored geometry for good.
extern "C" int mjw_set_media_nodal(mjw_problem *p, const double *rho_q, const double *mu_q) {
  if (p && p->pml) FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_media_nodal: not available with PML layers");
  if (!p || !rho_q) FAIL(p, MJW_ERR_ARG, "mjw_set_media_nodal: NULL");
  if (p->ncomp != 1) {
    // elastic: the density alone varies (it enters the mass only; a
    // varying stiffness tensor is not supported)
    if (mu_q)
      FAIL(p, MJW_ERR_UNSUPPORTED,
           "mjw_set_media_nodal: elastic problems take rho only (mu_q = NULL)");
    HIPCHK(p, hipSetDevice(p->device));
    const int64_t E = p->E, nd = p->nd;
    std::vector<double> rq((size_t)E * nd);
    for (int64_t es = 0; es < E; ++es)
      for (int64_t q = 0; q < nd; ++q) {
        const double r = rho_q[p->perm[es] * nd + q];
        if (!(r > 0) || !std::isfinite(r))
          FAIL(p, MJW_ERR_ARG, "mjw_set_media_nodal: rho must be finite, > 0");
        rq[(size_t)es * nd + q] = r;
      }
    std::vector<double> ms;
    const int rc = lumped_integral(p, rq.data(), true, ms);
    if (rc) return rc;
    p->mass = ms;  // solver order; finish_mass renumbers and uploads it
    HIPCHK(p, finish_mass(p));
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    p->inited = false;
    return MJW_OK;
  }
  if (!mu_q) FAIL(p, MJW_ERR_ARG, "mjw_set_media_nodal: NULL mu_q");
  HIPCHK(p, hipSetDevice(p->device));
  const int64_t E = p->E, nd = p->nd, N = p->N;
  std::vector<double> rq((size_t)E * nd), mq((size_t)E * nd);
  for (int64_t es = 0; es < E; ++es) {
    const int64_t e = p->perm[es];
    for (int64_t q = 0; q < nd; ++q) {
      const double r = rho_q[e * nd + q], m = mu_q[e * nd + q];
      if (!(r > 0) || !(m > 0) || !std::isfinite(r) || !std::isfinite(m))
        FAIL(p, MJW_ERR_ARG, "mjw_set_media_nodal: rho and mu must be finite, > 0");
      rq[(size_t)es * nd + q] = r;
      mq[(size_t)es * nd + q] = m;
    }
  }
  double xg[kMaxN], wg[kMaxN], Dg[kMaxN * kMaxN];
  gl_rule(p->order, xg, wg);
  gl_deriv(p->order, xg, Dg);
  std::vector<double> xyz_s((size_t)N * p->dim);
  for (int64_t g = 0; g < N; ++g)
    for (int a = 0; a < p->dim; ++a) xyz_s[(size_t)p->nperm[g] * p->dim + a] = p->xyz[g * p->dim + a];
  double *d_xyz = nullptr, *d_Dg = nullptr, *d_w = nullptr, *d_Mloc = nullptr, *d_rq = nullptr,
         *d_mq = nullptr;
  int *d_bad = nullptr;
  hipStream_t st = p->stream;
  const size_t bq = (size_t)E * nd * 8;
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t x) { if (e == hipSuccess) e = x; return e == hipSuccess; };
  ok(hipMalloc(&d_xyz, xyz_s.size() * 8));
  ok(hipMalloc(&d_Dg, p->n * p->n * 8));
  ok(hipMalloc(&d_w, p->n * 8));
  ok(hipMalloc(&d_Mloc, bq));
  ok(hipMalloc(&d_rq, bq));
  ok(hipMalloc(&d_mq, bq));
  if (!p->d_G) ok(dalloc(p, &p->d_G, (size_t)E * nd * (p->dim == 2 ? 3 : 6) * (p->geom == kG64 ? 8 : 4)));
  ok(hipMalloc(&d_bad, 4));
  ok(hipMemsetAsync(d_bad, 0, 4, st));
  ok(hipMemcpyAsync(d_xyz, xyz_s.data(), xyz_s.size() * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_Dg, Dg, p->n * p->n * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_w, wg, p->n * 8, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_rq, rq.data(), bq, hipMemcpyHostToDevice, st));
  ok(hipMemcpyAsync(d_mq, mq.data(), bq, hipMemcpyHostToDevice, st));
  if (e == hipSuccess &&
      launch_geometry_full(p->dim, p->order, 1, E, p->d_map, d_xyz, d_Dg, d_w, p->d_rho, p->d_mu,
                           d_rq, d_mq, p->geom == kG64 ? (double *)p->d_G : nullptr,
                           p->geom == kG32 ? (float *)p->d_G : nullptr, d_Mloc, d_bad, st))
    e = hipErrorLaunchFailure;
  for (int c = 0; c < p->ncolor && e == hipSuccess; ++c)
    if (launch_mass_assemble(p->d_map, d_Mloc, E, p->dim, p->order, p->coff[c], p->coff[c + 1],
                             p->d_M, st))
      e = hipErrorLaunchFailure;
  int bad = 0;
  ok(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, st));
  ok(hipMemcpyAsync(p->mass.data(), p->d_M, N * 8, hipMemcpyDeviceToHost, st));
  ok(hipStreamSynchronize(st));
  dfree(d_xyz); dfree(d_Dg); dfree(d_w); dfree(d_Mloc); dfree(d_rq); dfree(d_mq); dfree(d_bad);
  if (e != hipSuccess) FAIL(p, MJW_ERR_HIP, std::string("mjw_set_media_nodal: ") + hipGetErrorString(e));
  if (bad) FAIL(p, MJW_ERR_GEOMETRY, "non-positive Jacobian: degenerate or inverted element");
  HIPCHK(p, finish_mass(p));
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->bexec) { (void)hipGraphExecDestroy(p->bexec); p->bexec = nullptr; }
  p->tri_ok = false;
  p->geo = kGeoStored;
  p->nodal_media = true;
  p->inited = false;
  p->adj_ready = false;
  return MJW_OK;
}

extern "C" int mjw_adjoint_forward(mjw_problem *p, int nsteps, int nsnap, double *traces) {
  if (p && p->pml) FAIL(p, MJW_ERR_UNSUPPORTED, "adjoint: not available with PML layers");
  if (p && p->torder == 4)
    FAIL(p, MJW_ERR_UNSUPPORTED, "adjoint: the second-order scheme only");
  if (!p || nsteps < 2 || nsnap < 1 || !traces) FAIL(p, MJW_ERR_ARG, "mjw_adjoint_forward: bad args");
  if (p->ncomp != 1 || p->nshots != 1)
    FAIL(p, MJW_ERR_UNSUPPORTED, "adjoint: single-shot acoustic problems only");
  if (p->nodal_media)
    FAIL(p, MJW_ERR_UNSUPPORTED, "adjoint: per-element mu only (the problem has nodal media)");
  if (!p->damp_S.empty()) FAIL(p, MJW_ERR_UNSUPPORTED, "adjoint: undamped problems only");
  if (!p->inited || p->step != 0 || !p->zero_ic)
    FAIL(p, MJW_ERR_STATE, "adjoint: call mjw_leapfrog_init with zero initial state first");
  if (p->nrec == 0) FAIL(p, MJW_ERR_STATE, "adjoint: no receivers");
  HIPCHK(p, hipSetDevice(p->device));
  for (void *q : p->snaps) dfree(q);
  p->snaps.clear();
  p->snap_step.clear();
  const int64_t N1 = nsteps - 1;                      // steps run: U^0..U^N
  const int64_t K = (N1 + nsnap - 1) / nsnap;         // segment length
  const size_t vb = (size_t)p->N * ssize(p);
  hipEvent_t e0, e1;
  HIPCHK(p, hipEventCreate(&e0));
  HIPCHK(p, hipEventCreate(&e1));
  HIPCHK(p, hipEventRecord(e0, p->stream));
  for (int64_t k = 0; k < N1; k += K) {
    void *s = nullptr;
    HIPCHK(p, hipMalloc(&s, 2 * vb));
    p->snaps.push_back(s);
    p->snap_step.push_back(k);
    HIPCHK(p, hipMemcpyAsync(s, p->d_U[p->cur], vb, hipMemcpyDeviceToDevice, p->stream));
    HIPCHK(p, hipMemcpyAsync((char *)s + vb, p->d_U[p->cur ^ 1], vb, hipMemcpyDeviceToDevice,
                             p->stream));
    const int64_t m = std::min(K, N1 - k);
    int rc = mjw_leapfrog_advance(p, (int)m, traces + k * p->nrec, 1);
    if (rc) return rc;
  }
  {  // row N-1 = R U^{N-1} (current iterate)
    int rc = mjw_record(p, traces + N1 * p->nrec);
    if (rc) return rc;
  }
  {  // row N = R U^N (the next iterate, cur^1)
    KCHK(p, ensure_traces(p, 1));
    int64_t ctl[2] = {0, 0};
    HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, p->stream));
    TailArgs t{};
    t.state = p->state; t.ctl = p->d_ctl; t.nrec = p->nrec; t.nd = p->nd; t.rec_gid = p->d_rec_gid;
    t.rec_w = p->d_rec_w; t.rec_comp = p->d_rec_comp; t.U = p->d_U[p->cur ^ 1]; t.P = nullptr;
    t.traces = p->d_traces; t.N = p->N; t.nshots = 1; t.advance = 0; t.row_offset = 0;
    KCHK(p, launch_tail_n(t, p->ncomp, p->stream));
    HIPCHK(p, hipMemcpyAsync(traces + (size_t)nsteps * p->nrec, p->d_traces, (size_t)p->nrec * 8,
                             hipMemcpyDeviceToHost, p->stream));
  }
  HIPCHK(p, hipEventRecord(e1, p->stream));
  HIPCHK(p, hipEventSynchronize(e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->ms_adj_fwd = ms;
  p->adj_nsteps = nsteps;
  p->adj_ready = true;
  return MJW_OK;
}

// one backward step j: buffers cur (U^j, nu_j) and other (U^{j+1}, nu_{j+1})
static int enqueue_bwd_step(mjw_problem *p, void *cur, void *other) {
  for (int c = 0; c < p->ncolor; ++c) {
    StiffArgs a = stiff_args(p, kLeap, cur, other, 2);
    a.e0 = p->coff[c];
    a.e1 = p->coff[c + 1];
    a.grad = p->d_grad;
    int rc = launch_stiffness(a, p->stream);
    if (rc) return rc;
  }
  TailArgs t{};
  t.state = p->state; t.ctl = p->d_ctl; t.pulse = p->d_pulse; t.npulse = p->npulse;
  t.nsrc = p->nsrc; t.src_dof = p->d_src_dof; t.src_shot = p->d_src_shot; t.src_val = p->d_src_val;
  t.sdiag = p->d_sdiag; t.nrec = 0; t.nd = p->nd; t.U = cur; t.P = other; t.traces = nullptr;
  t.N = p->N; t.nshots = 2; t.advance = -1; t.row_offset = 0;
  t.nadj = p->nadj; t.adj_node = p->d_adj_node; t.adj_ptr = p->d_adj_ptr; t.adj_rec = p->d_adj_rec;
  t.adj_w = p->d_adj_w; t.adj_g = p->d_adjg; t.adj_row = 1; t.adj_nrec = p->nrec; t.adj_shot = 1;
  return launch_tail_n(t, p->ncomp, p->stream);
}

extern "C" int mjw_adjoint_backward(mjw_problem *p, const double *dJ, double *grad) {
  if (!p || !dJ || !grad) FAIL(p, MJW_ERR_ARG, "mjw_adjoint_backward: NULL");
  if (!p->adj_ready) FAIL(p, MJW_ERR_STATE, "adjoint: run mjw_adjoint_forward first");
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  const int64_t Nn = p->adj_nsteps;                    // traces rows 0..Nn
  const size_t vb = (size_t)p->N * ssize(p);
  const int64_t rows = Nn + 1;
  if (rows * p->nrec > p->adjg_cap) {
    dfree(p->d_adjg);
    p->d_adjg = nullptr;
    HIPCHK(p, hipMalloc(&p->d_adjg, rows * p->nrec * 8));
    p->adjg_cap = rows * p->nrec;
    if (p->bexec) { (void)hipGraphExecDestroy(p->bexec); p->bexec = nullptr; }
  }
  for (int b = 0; b < 2; ++b)
    if (!p->d_pair[b]) {
      HIPCHK(p, dalloc(p, &p->d_pair[b], 2 * vb));
      if (p->bexec) { (void)hipGraphExecDestroy(p->bexec); p->bexec = nullptr; }
    }
  if (!p->d_grad) HIPCHK(p, dalloc(p, &p->d_grad, p->E * 8));
  HIPCHK(p, hipMemcpyAsync(p->d_adjg, dJ, rows * p->nrec * 8, hipMemcpyHostToDevice, st));
  HIPCHK(p, hipMemsetAsync(p->d_grad, 0, p->E * 8, st));
  hipEvent_t e0, e1;
  HIPCHK(p, hipEventCreate(&e0));
  HIPCHK(p, hipEventCreate(&e1));
  HIPCHK(p, hipEventRecord(e0, st));
  // cur = (U^{N-1}, nu_{N-1} = S h^N), other = (U^N, nu_N = 0)
  void *cur = p->d_pair[0], *oth = p->d_pair[1];
  HIPCHK(p, hipMemcpyAsync(cur, p->d_U[p->cur], vb, hipMemcpyDeviceToDevice, st));
  HIPCHK(p, hipMemcpyAsync(oth, p->d_U[p->cur ^ 1], vb, hipMemcpyDeviceToDevice, st));
  HIPCHK(p, hipMemsetAsync((char *)cur + vb, 0, vb, st));
  HIPCHK(p, hipMemsetAsync((char *)oth + vb, 0, vb, st));
  {
    int64_t ctl[2] = {Nn - 1, 0};   // adj row = step + 1 = N
    HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, st));
    TailArgs t{};
    t.state = p->state; t.ctl = p->d_ctl; t.sdiag = p->d_sdiag; t.U = cur; t.P = cur;
    t.N = p->N; t.nshots = 2; t.advance = 0;
    t.nadj = p->nadj; t.adj_node = p->d_adj_node; t.adj_ptr = p->d_adj_ptr;
    t.adj_rec = p->d_adj_rec; t.adj_w = p->d_adj_w; t.adj_g = p->d_adjg; t.adj_row = 1;
    t.adj_nrec = p->nrec; t.adj_shot = 1;
    KCHK(p, launch_tail_n(t, p->ncomp, st));
  }
  // step counter = j - 1 (pulse index of the reverse step), decremented by the tail
  {
    int64_t ctl[2] = {Nn - 2, 0};
    HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, st));
  }
  // re-anchoring points: after the step j = k_s + 1 the current pair holds
  // (U^{k_s}, U^{k_s+1})
  std::vector<int> anchor(Nn + 1, -1);
  for (size_t s = 0; s < p->snap_step.size(); ++s) anchor[p->snap_step[s]] = (int)s;
  int64_t j = Nn - 1;
  while (j >= 1) {
    // run backward steps until the next anchor (k_s = j - 1)
    int64_t stop = j;  // last j of this run
    while (stop >= 1 && anchor[stop - 1] < 0) --stop;
    // steps j, j-1, ..., stop (stop = k_s + 1 or 1)
    if (stop < 1) stop = 1;
    int64_t cnt = j - stop + 1;
    bool parity0 = true;  // cur == d_pair[0]
    (void)parity0;
    // two-step graph when the buffers are in their canonical order
    while (cnt >= 2 && cur == p->d_pair[0]) {
      if (!p->bexec) {
        hipGraph_t g = nullptr;
        HIPCHK(p, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_bwd_step(p, p->d_pair[0], p->d_pair[1]);
        if (!rc) rc = enqueue_bwd_step(p, p->d_pair[1], p->d_pair[0]);
        hipError_t ee = hipStreamEndCapture(st, &g);
        if (rc || ee != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          FAIL(p, MJW_ERR_HIP, "adjoint graph capture failed");
        }
        ee = hipGraphInstantiate(&p->bexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ee != hipSuccess) { p->bexec = nullptr; FAIL(p, MJW_ERR_HIP, "graph instantiate"); }
      }
      HIPCHK(p, hipGraphLaunch(p->bexec, st));
      cnt -= 2;
      j -= 2;
    }
    while (cnt > 0) {
      KCHK(p, enqueue_bwd_step(p, cur, oth));
      std::swap(cur, oth);
      --cnt;
      --j;
    }
    if (j >= 0 && anchor[j] >= 0 && j < Nn - 1) {
      // cur holds (U^j, nu_j), oth holds (U^{j+1}, nu_{j+1}): restore U exactly
      void *s = p->snaps[anchor[j]];
      HIPCHK(p, hipMemcpyAsync(cur, s, vb, hipMemcpyDeviceToDevice, st));
      HIPCHK(p, hipMemcpyAsync(oth, (char *)s + vb, vb, hipMemcpyDeviceToDevice, st));
    }
  }
  HIPCHK(p, hipEventRecord(e1, st));
  HIPCHK(p, hipEventSynchronize(e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->ms_adj_bwd = ms;
  std::vector<double> gs(p->E);
  HIPCHK(p, hipMemcpy(gs.data(), p->d_grad, p->E * 8, hipMemcpyDeviceToHost));
  for (int64_t es = 0; es < p->E; ++es) grad[p->perm[es]] = gs[es];
  return MJW_OK;
}

extern "C" int mjw_adjoint_timing(const mjw_problem *p, double *ms2) {
  if (!p || !ms2) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL");
  ms2[0] = p->ms_adj_fwd;
  ms2[1] = p->ms_adj_bwd;
  return MJW_OK;
}

// ------------------------------------------------------------------ batch
// Independent 3-D acoustic problems (uneven sizes and orders) advanced
// together: per step one stiffness launch per colour index over the work
// items of every problem having that colour (k_batch3d dispatches on the
// element order), then one tail block per problem; the step loop is captured
// in a HIP graph.
// ---------------------------------------------------------- PML layers
// Montjoie's first-order split system (ModifiedEquationSystemIterator, order
// 2, Instationary/ModifiedEquation.cxx:511-560,773-812) with the element
// matrices of the first-order Helmholtz formulation (Elliptic/Helmholtz/
// VarHelmholtz.cxx:133-230,296-378,620-660,3277-3720) and the block-diagonal
// mass / damping operators of Hyperbolic/Acoustic/TimeAcoustic.cxx:1079-1235.

// 2x2 / 3x3 inverse (adjugate over determinant)
static void inv_small(int d, const double *P, double *Q) {
  if (d == 1) { Q[0] = 1.0 / P[0]; return; }
  if (d == 2) {
    const double det = P[0] * P[3] - P[1] * P[2];
    Q[0] = P[3] / det; Q[1] = -P[1] / det; Q[2] = -P[2] / det; Q[3] = P[0] / det;
    return;
  }
  const double c00 = P[4] * P[8] - P[5] * P[7], c01 = P[5] * P[6] - P[3] * P[8],
               c02 = P[3] * P[7] - P[4] * P[6];
  const double det = P[0] * c00 + P[1] * c01 + P[2] * c02;
  Q[0] = c00 / det; Q[1] = (P[2] * P[7] - P[1] * P[8]) / det; Q[2] = (P[1] * P[5] - P[2] * P[4]) / det;
  Q[3] = c01 / det; Q[4] = (P[0] * P[8] - P[2] * P[6]) / det; Q[5] = (P[2] * P[3] - P[0] * P[5]) / det;
  Q[6] = c02 / det; Q[7] = (P[1] * P[6] - P[0] * P[7]) / det; Q[8] = (P[0] * P[4] - P[1] * P[3]) / det;
}

extern "C" int mjw_set_pml(mjw_problem *p, const uint8_t *pml, const double *tau) {
  if (!p || !pml || !tau) FAIL(p, MJW_ERR_ARG, "mjw_set_pml: NULL");
  if (p->ncomp != 1 || p->nshots != 1)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_pml: single-shot acoustic problems only");
  if (p->torder == 4)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_pml: the first-order system is of order two in time");
  if (!p->damp_S.empty())
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_pml: physical damping with PML layers is not supported "
                                 "(VarHelmholtz.cxx:411-422 refuses it inside the layers)");
  if (p->nodal_media)
    FAIL(p, MJW_ERR_UNSUPPORTED, "mjw_set_pml: per-element media only");
  if (p->pml) FAIL(p, MJW_ERR_STATE, "mjw_set_pml: the layers are already set");
  HIPCHK(p, hipSetDevice(p->device));
  HIPCHK(p, hipStreamSynchronize(p->stream));
  const int d = p->dim, n = p->n, nd = p->nd, P = p->P;
  const int64_t E = p->E, N = p->N, EN = E * nd;
  for (int64_t i = 0; i < EN * d; ++i)
    if (!(tau[i] >= 0.0) || !std::isfinite(tau[i]))
      FAIL(p, MJW_ERR_ARG, "mjw_set_pml: tau must be finite and >= 0");
  double xg[kMaxN], wg[kMaxN], Dg[kMaxN * kMaxN];
  gl_rule(p->order, xg, wg);
  gl_deriv(p->order, xg, Dg);
  std::vector<double> hG((size_t)d * d * EN), hB(EN), ht((size_t)d * EN, 0.0);
  std::vector<uint8_t> hflag(E), blk(N, 0);
  std::vector<double> Dc((size_t)d * N, 0.0);
  std::vector<int32_t> bidx(N, -1);
  std::vector<int32_t> pnode;
  std::vector<double> Sb;
  std::vector<double> rho_e(E), mu_e(E);
  HIPCHK(p, hipMemcpy(rho_e.data(), p->d_rho, E * 8, hipMemcpyDeviceToHost));
  HIPCHK(p, hipMemcpy(mu_e.data(), p->d_mu, E * 8, hipMemcpyDeviceToHost));
  for (int64_t es = 0; es < E; ++es) {
    const int64_t e = p->perm[es];
    hflag[es] = pml[e] ? 1 : 0;
    const int32_t *en = &p->en[e * nd];
    for (int q = 0; q < nd; ++q) {
      const int i = q % n, j = (q / n) % n, k = q / P;
      const int ix[3] = {i, j, k};
      // DF[a][b] = d x_a / d xi_b through the GL derivative matrix
      double DF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int b = 0; b < d; ++b)
        for (int m = 0; m < n; ++m) {
          int jx[3] = {i, j, k};
          jx[b] = m;
          const int qm = jx[0] + n * jx[1] + P * jx[2];
          const double c = Dg[ix[b] * n + m];
          for (int a = 0; a < d; ++a) DF[a * d + b] += c * p->xyz[(int64_t)en[qm] * d + a];
        }
      double adj[9], J;
      if (d == 2) {
        J = DF[0] * DF[3] - DF[1] * DF[2];
        adj[0] = DF[3]; adj[1] = -DF[1]; adj[2] = -DF[2]; adj[3] = DF[0];
      } else {
        adj[0] = DF[4] * DF[8] - DF[5] * DF[7]; adj[1] = DF[2] * DF[7] - DF[1] * DF[8];
        adj[2] = DF[1] * DF[5] - DF[2] * DF[4]; adj[3] = DF[5] * DF[6] - DF[3] * DF[8];
        adj[4] = DF[0] * DF[8] - DF[2] * DF[6]; adj[5] = DF[2] * DF[3] - DF[0] * DF[5];
        adj[6] = DF[3] * DF[7] - DF[4] * DF[6]; adj[7] = DF[1] * DF[6] - DF[0] * DF[7];
        adj[8] = DF[0] * DF[4] - DF[1] * DF[3];
        J = DF[0] * adj[0] + DF[1] * adj[3] + DF[2] * adj[6];
      }
      if (!(J > 0.0)) FAIL(p, MJW_ERR_GEOMETRY, "mjw_set_pml: degenerate element");
      double w = 1.0;
      for (int a = 0; a < d; ++a) w *= wg[ix[a]];
      const int64_t pidx = ((int64_t)k * E + es) * P + (q % P);
      for (int a = 0; a < d * d; ++a) hG[(size_t)a * EN + pidx] = w * adj[a];
      const double Jw = J * w;
      hB[pidx] = Jw / mu_e[es];
      if (!pml[e]) continue;
      const double *t = tau + (e * nd + q) * d;
      for (int a = 0; a < d; ++a) ht[(size_t)a * EN + pidx] = t[a];
      const int32_t g = p->nperm[en[q]];
      if (bidx[g] < 0) {
        bidx[g] = (int32_t)pnode.size();
        pnode.push_back(g);
        Sb.resize(Sb.size() + d * d, 0.0);
      }
      blk[g] = 1;
      const double m = rho_e[es] * Jw;
      for (int c = 1; c < d; ++c) Dc[(size_t)c * N + g] += m;
      double *S = &Sb[(size_t)bidx[g] * d * d];
      if (d == 2) {
        const double sp = rho_e[es] * (0.5 * (t[0] + t[1])) * Jw;
        const double sm = rho_e[es] * (0.5 * (t[0] - t[1])) * Jw;
        S[0] += sp; S[3] += sp; S[1] += sm; S[2] += sm;
      } else {
        const double s0 = rho_e[es] * (t[0] + t[1] + t[2]) * Jw / 3.0;
        const double d1 = rho_e[es] * (t[0] - t[1]) * Jw / 3.0;
        const double d2 = rho_e[es] * (t[0] - t[2]) * Jw / 3.0;
        const double d3 = rho_e[es] * (t[1] - t[2]) * Jw / 3.0;
        S[0] += s0; S[1] += d1 - d3; S[2] += d2 + d3;
        S[3] += d1; S[4] += s0 + d3; S[5] += d1;
        S[6] += d2; S[7] += d2; S[8] += s0 - d3;
      }
    }
  }
  for (int64_t g = 0; g < N; ++g) Dc[p->nperm[g]] = p->mass[g];  // Dirichlet rows 1
  for (size_t t = 0; t < pnode.size(); ++t) {
    const int32_t g = pnode[t];
    if (p->dir[p->ninv[g]])
      for (int c = 0; c < d; ++c) { Sb[t * d * d + c] = 0.0; Sb[t * d * d + c * d] = 0.0; }
  }
  p->pml_D = Dc;
  p->pml_S = Sb;
  p->pml_node = pnode;
  // device arrays
  const size_t sz = ssize(p);
  HIPCHK(p, dalloc(p, &p->d_pml_G, hG.size() * 8));
  HIPCHK(p, dalloc(p, &p->d_pml_B, hB.size() * 8));
  HIPCHK(p, dalloc(p, &p->d_pml_tau, ht.size() * 8));
  HIPCHK(p, dalloc(p, &p->d_pml_Dm, (size_t)n * n * 8));
  HIPCHK(p, dalloc(p, &p->d_pml_flag, (size_t)E));
  HIPCHK(p, dalloc(p, &p->d_pml_blk, (size_t)N));
  HIPCHK(p, dalloc(p, &p->d_pml_node, std::max<size_t>(pnode.size(), 1) * 4));
  HIPCHK(p, dalloc(p, &p->d_pml_V, (size_t)d * EN * sz));
  HIPCHK(p, dalloc(p, &p->d_pml_Pr, (size_t)d * N * sz));
  HIPCHK(p, dalloc(p, &p->d_pml_sa, (size_t)N * 8));
  HIPCHK(p, dalloc(p, &p->d_pml_sinv, (size_t)N * 8));
  HIPCHK(p, dalloc(p, &p->d_pml_A, std::max<size_t>(Sb.size(), 1) * 8));
  HIPCHK(p, dalloc(p, &p->d_pml_Pinv, std::max<size_t>(Sb.size(), 1) * 8));
  HIPCHK(p, hipMemcpy(p->d_pml_G, hG.data(), hG.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(p, hipMemcpy(p->d_pml_B, hB.data(), hB.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(p, hipMemcpy(p->d_pml_tau, ht.data(), ht.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(p, hipMemcpy(p->d_pml_Dm, Dg, (size_t)n * n * 8, hipMemcpyHostToDevice));
  HIPCHK(p, hipMemcpy(p->d_pml_flag, hflag.data(), (size_t)E, hipMemcpyHostToDevice));
  HIPCHK(p, hipMemcpy(p->d_pml_blk, blk.data(), (size_t)N, hipMemcpyHostToDevice));
  if (!pnode.empty())
    HIPCHK(p, hipMemcpy(p->d_pml_node, pnode.data(), pnode.size() * 4, hipMemcpyHostToDevice));
  // the state: u and its split copies in one buffer, u first
  void *U = nullptr;
  HIPCHK(p, dalloc(p, &U, (size_t)d * N * sz));
  dfree(p->d_U[0]);
  p->d_U[0] = U;
  p->cur = 0;
  p->pml = true;
  p->inited = false;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  return MJW_OK;
}

static int pml_init(mjw_problem *p, double t_begin, double dt, const double *u0,
                    const double *du0) {
  if (u0 || du0)
    FAIL(p, MJW_ERR_UNSUPPORTED, "PML layers: the first-order system starts from zero data "
                                 "(u0 = du0 = NULL)");
  HIPCHK(p, hipSetDevice(p->device));
  hipStream_t st = p->stream;
  const int d = p->dim;
  const int64_t N = p->N, EN = p->E * p->nd;
  p->dt = dt;
  p->t_begin = t_begin;
  p->zero_ic = true;
  p->adj_ready = false;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }  // dt is baked in
  std::vector<double> sa(N), sinv(N);
  for (int64_t i = 0; i < N; ++i) {
    sa[i] = p->pml_D[i];
    sinv[i] = p->dir[p->ninv[i]] ? 0.0 : 1.0 / p->pml_D[i];
  }
  const size_t nb = p->pml_node.size();
  std::vector<double> A(nb * d * d), Pinv(nb * d * d);
  for (size_t t = 0; t < nb; ++t) {
    const int32_t g = p->pml_node[t];
    double Pm[9];
    for (int r = 0; r < d; ++r)
      for (int c = 0; c < d; ++c) {
        const double dg = r == c ? p->pml_D[(size_t)r * N + g] : 0.0;
        const double sv = p->pml_S[(t * d + r) * d + c];
        A[(t * d + r) * d + c] = dg - 0.5 * dt * sv;
        Pm[r * d + c] = dg + 0.5 * dt * sv;
      }
    inv_small(d, Pm, &Pinv[t * d * d]);
  }
  HIPCHK(p, hipMemcpy(p->d_pml_sa, sa.data(), N * 8, hipMemcpyHostToDevice));
  HIPCHK(p, hipMemcpy(p->d_pml_sinv, sinv.data(), N * 8, hipMemcpyHostToDevice));
  if (nb) {
    HIPCHK(p, hipMemcpy(p->d_pml_A, A.data(), A.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(p, hipMemcpy(p->d_pml_Pinv, Pinv.data(), Pinv.size() * 8, hipMemcpyHostToDevice));
  }
  const size_t sz = ssize(p);
  // U^0 = 0, V^{1/2} = 0 (zero data), P = R^S V^{1/2} = 0
  HIPCHK(p, hipMemsetAsync(p->d_U[0], 0, (size_t)d * N * sz, st));
  HIPCHK(p, hipMemsetAsync(p->d_pml_V, 0, (size_t)d * EN * sz, st));
  HIPCHK(p, hipMemsetAsync(p->d_pml_Pr, 0, (size_t)d * N * sz, st));
  p->cur = 0;
  p->step = 0;
  int64_t ctl[2] = {0, 0};
  HIPCHK(p, hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, st));
  HIPCHK(p, hipStreamSynchronize(st));
  p->inited = true;
  return MJW_OK;
}

// one step of the first-order system: traces of u^n and F_n b into P, the
// node update to U^{n+1}, then per colour V^{n+3/2} and P = R^S V^{n+3/2}
static int enqueue_step_pml(mjw_problem *p, double *traces) {
  TailArgs t{};
  t.state = p->state; t.ctl = p->d_ctl; t.pulse = p->d_pulse; t.npulse = p->npulse;
  t.nsrc = p->nsrc; t.src_dof = p->d_src_dof; t.src_shot = p->d_src_shot; t.src_val = p->d_src_val;
  t.sdiag = nullptr; t.nrec = traces ? p->nrec : 0; t.nd = p->nd; t.rec_gid = p->d_rec_gid;
  t.rec_w = p->d_rec_w; t.rec_comp = p->d_rec_comp; t.U = p->d_U[0]; t.P = p->d_pml_Pr;
  t.traces = (traces && p->nrec) ? traces : nullptr; t.N = p->N; t.nshots = 1; t.ilv = 0;
  t.advance = 1;
  t.row_offset = 0;
  int rc = launch_tail_n(t, 1, p->stream);
  if (rc) return rc;
  PmlArgs a{};
  a.dim = p->dim; a.order = p->order; a.state = p->state; a.E = p->E; a.N = p->N;
  a.map = p->d_map; a.G = p->d_pml_G; a.B = p->d_pml_B; a.tau = p->d_pml_tau;
  a.pml = p->d_pml_flag; a.Dm = p->d_pml_Dm; a.U = p->d_U[0]; a.V = p->d_pml_V;
  a.Pr = p->d_pml_Pr; a.dt = p->dt; a.npn = (int64_t)p->pml_node.size(); a.pnode = p->d_pml_node;
  a.A = p->d_pml_A; a.Pinv = p->d_pml_Pinv; a.dir = p->d_dir; a.blk = p->d_pml_blk;
  a.sa = p->d_pml_sa; a.sinv = p->d_pml_sinv;
  rc = launch_pml_nodes(a, p->stream);
  if (rc) return rc;
  for (int c = 0; c < p->ncolor; ++c) {
    a.e0 = p->coff[c];
    a.e1 = p->coff[c + 1];
    rc = launch_pml_elem(a, p->stream);
    if (rc) return rc;
  }
  return 0;
}

struct mjw_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<mjw_problem *> probs;
  int state = kF64, geom = kG64;
  int ncolor = 0;
  // work order: 0 items by decreasing order, 1 problem-major (one launch per
  // colour over every order); 2, 3, 4 one launch per (colour, order) with
  // blocks of about 256, 512, 128 threads
  int order_mode = 0;
  BatchProb *d_probs = nullptr;
  std::vector<BatchItem *> d_items;   // per colour
  std::vector<int> nitems;
  // grouped modes: per colour, (order, first item, item count) of each order
  std::vector<std::vector<std::array<int, 3>>> segs;
  TailArgs *d_tail[2] = {nullptr, nullptr};
  hipGraphExec_t gexec = nullptr;
  bool g_traces = false;
  int64_t last_nsteps = 0;
  double ms_step = 0.0;
  std::string err;
};

static void batch_free_items(mjw_batch *b) {
  for (BatchItem *q : b->d_items) dfree(q);
  b->d_items.clear();
  b->nitems.clear();
  b->segs.clear();
}

static int batch_build_items(mjw_batch *b) {
  batch_free_items(b);
  b->ncolor = 0;
  for (mjw_problem *p : b->probs) b->ncolor = std::max(b->ncolor, p->ncolor);
  for (int c = 0; c < b->ncolor; ++c) {
    std::vector<std::pair<int, BatchItem>> items;  // (order key, item)
    for (size_t i = 0; i < b->probs.size(); ++i) {
      mjw_problem *p = b->probs[i];
      if (c >= p->ncolor) continue;
      const int epb =
          b->order_mode >= 2 ? batch_group_epb(p->order, b->order_mode - 2) : batch_epb(p->order);
      for (int64_t eb = p->coff[c]; eb < p->coff[c + 1]; eb += epb) {
        BatchItem it{};
        it.prob = (int32_t)i;
        it.eb = eb;
        it.e1 = p->coff[c + 1];
        items.push_back({b->order_mode == 1 ? (int)i : -p->order, it});
      }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const auto &x, const auto &y) { return x.first < y.first; });
    std::vector<BatchItem> h(items.size());
    for (size_t k = 0; k < items.size(); ++k) h[k] = items[k].second;
    std::vector<std::array<int, 3>> sg;  // runs of one order (items sorted by order)
    for (size_t k = 0; k < items.size(); ++k) {
      const int ord = -items[k].first;
      if (sg.empty() || sg.back()[0] != ord) sg.push_back({ord, (int)k, 0});
      sg.back()[2]++;
    }
    b->segs.push_back(sg);
    BatchItem *d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(BatchItem)) != hipSuccess) return -1;
    if (!h.empty() &&
        hipMemcpy(d, h.data(), h.size() * sizeof(BatchItem), hipMemcpyHostToDevice) != hipSuccess)
      return -1;
    b->d_items.push_back(d);
    b->nitems.push_back((int)h.size());
  }
  return 0;
}

#define BFAIL(b, code, msg)               \
  do {                                    \
    if (b) (b)->err = (msg);              \
    g_err = (msg);                        \
    return (code);                        \
  } while (0)

extern "C" int mjw_batch_create(mjw_batch **out, int nprob, mjw_problem *const *probs) {
  if (!out || nprob < 1 || !probs) BFAIL((mjw_batch *)nullptr, MJW_ERR_ARG, "mjw_batch_create: bad args");
  *out = nullptr;
  for (int i = 0; i < nprob; ++i) {
    const mjw_problem *p = probs[i];
    if (!p) BFAIL((mjw_batch *)nullptr, MJW_ERR_ARG, "mjw_batch_create: NULL problem");
    if (p->dim != 3 || p->ncomp != 1 || p->nshots != 1)
      BFAIL((mjw_batch *)nullptr, MJW_ERR_UNSUPPORTED, "batch: single-shot 3-D acoustic problems");
    if (!p->damp_S.empty())
      BFAIL((mjw_batch *)nullptr, MJW_ERR_UNSUPPORTED, "batch: undamped problems only");
    if (p->pml)
      BFAIL((mjw_batch *)nullptr, MJW_ERR_UNSUPPORTED, "batch: problems without PML layers only");
    if (p->torder == 4)
      BFAIL((mjw_batch *)nullptr, MJW_ERR_UNSUPPORTED, "batch: second-order scheme only");
    if (p->device != probs[0]->device || p->state != probs[0]->state ||
        p->geom != probs[0]->geom)
      BFAIL((mjw_batch *)nullptr, MJW_ERR_ARG, "batch: problems must share device and precision");
  }
  mjw_batch *b = new mjw_batch();
  b->device = probs[0]->device;
  b->state = probs[0]->state;
  b->geom = probs[0]->geom;
  b->probs.assign(probs, probs + nprob);
  if (hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    BFAIL((mjw_batch *)nullptr, MJW_ERR_HIP, "batch: stream creation failed");
  }
  std::vector<BatchProb> hp(nprob);
  for (int i = 0; i < nprob; ++i) {
    mjw_problem *p = probs[i];
    BatchProb &d = hp[i];
    d.map = p->d_map; d.G = p->d_G; d.Dm = p->d_D; d.sdiag = p->d_sdiag;
    d.buf[0] = p->d_U[0]; d.buf[1] = p->d_U[1];
    d.E = p->E; d.N = p->N; d.order = p->order;
    d.geo = p->geo; d.coef = p->d_coef; d.glxw = p->d_glxw;
  }
  int rc = (int)hipMalloc(&b->d_probs, nprob * sizeof(BatchProb));
  if (!rc) rc = (int)hipMemcpy(b->d_probs, hp.data(), nprob * sizeof(BatchProb), hipMemcpyHostToDevice);
  if (!rc) rc = (int)hipMalloc(&b->d_tail[0], nprob * sizeof(TailArgs));
  if (!rc) rc = (int)hipMalloc(&b->d_tail[1], nprob * sizeof(TailArgs));
  if (!rc) rc = batch_build_items(b);
  if (rc) {
    mjw_batch_destroy(b);
    BFAIL((mjw_batch *)nullptr, MJW_ERR_HIP, "batch: setup failed");
  }
  *out = b;
  return MJW_OK;
}

extern "C" int mjw_batch_config(mjw_batch *b, int nxcd, int order_mode) {
  if (!b || order_mode < 0 || order_mode > 4) BFAIL(b, MJW_ERR_ARG, "mjw_batch_config: bad args");
  (void)nxcd;  // block order of the batch launches follows the work list
  if (order_mode != b->order_mode) {
    b->order_mode = order_mode;
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (batch_build_items(b)) BFAIL(b, MJW_ERR_HIP, "batch: work list rebuild failed");
  }
  return MJW_OK;
}

static int batch_enqueue_step(mjw_batch *b, int parity) {
  for (int c = 0; c < b->ncolor; ++c) {
    if (b->order_mode >= 2) {
      for (const auto &sg : b->segs[c]) {
        int rc = launch_batch3d_order(b->state, b->geom, sg[0], b->order_mode - 2, b->d_probs,
                                      b->d_items[c] + sg[1], sg[2], parity, b->stream);
        if (rc) return rc;
      }
      continue;
    }
    int rc = launch_batch3d(b->state, b->geom, b->d_probs, b->d_items[c], b->nitems[c], parity,
                            b->stream);
    if (rc) return rc;
  }
  return launch_tail_batch(b->state, b->d_tail[parity], (int)b->probs.size(), b->stream);
}

extern "C" int mjw_batch_advance(mjw_batch *b, int nsteps, int use_graph) {
  if (!b || nsteps < 0) BFAIL(b, MJW_ERR_ARG, "mjw_batch_advance: bad args");
  if (nsteps == 0) return MJW_OK;
  if (hipSetDevice(b->device) != hipSuccess) BFAIL(b, MJW_ERR_HIP, "hipSetDevice");
  const int parity = b->probs[0]->cur;
  bool traces = false;
  for (mjw_problem *p : b->probs) {
    if (!p->inited) BFAIL(b, MJW_ERR_STATE, "batch: every problem needs mjw_leapfrog_init");
    if (p->cur != parity) BFAIL(b, MJW_ERR_STATE, "batch: problems must be at the same parity");
    traces = traces || p->nrec > 0;
  }
  // per-problem tail arguments for both parities (read by the kernels at run time)
  std::vector<TailArgs> ta[2];
  for (int par = 0; par < 2; ++par) ta[par].resize(b->probs.size());
  for (size_t i = 0; i < b->probs.size(); ++i) {
    mjw_problem *p = b->probs[i];
    if (hipStreamSynchronize(p->stream) != hipSuccess) BFAIL(b, MJW_ERR_HIP, "sync");
    if (p->nrec && ensure_traces(p, nsteps)) BFAIL(b, MJW_ERR_NOMEM, "batch: traces buffer");
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
    for (int par = 0; par < 2; ++par) {
      TailArgs &t = ta[par][i];
      t = TailArgs{};
      t.state = p->state; t.ctl = p->d_ctl; t.pulse = p->d_pulse; t.npulse = p->npulse;
      t.nsrc = p->nsrc; t.src_dof = p->d_src_dof; t.src_shot = p->d_src_shot;
      t.src_val = p->d_src_val; t.sdiag = p->d_sdiag; t.nrec = p->nrec; t.nd = p->nd;
      t.rec_gid = p->d_rec_gid; t.rec_w = p->d_rec_w; t.rec_comp = p->d_rec_comp;
      t.U = p->d_U[par ^ 1]; t.P = p->d_U[par]; t.traces = p->nrec ? p->d_traces : nullptr;
      t.N = p->N; t.nshots = 1; t.advance = 1; t.row_offset = 1;
    }
    int64_t ctl[2] = {p->step, p->step};
    if (hipMemcpyAsync(p->d_ctl, ctl, 16, hipMemcpyHostToDevice, b->stream) != hipSuccess)
      BFAIL(b, MJW_ERR_HIP, "ctl");
    if (p->nrec) {  // row 0 = U^n
      TailArgs r{};
      r.state = p->state; r.ctl = p->d_ctl; r.nrec = p->nrec; r.nd = p->nd;
      r.rec_gid = p->d_rec_gid; r.rec_w = p->d_rec_w; r.rec_comp = p->d_rec_comp;
      r.U = p->d_U[p->cur]; r.traces = p->d_traces; r.N = p->N; r.nshots = 1;
      if (launch_tail_n(r, 1, b->stream)) BFAIL(b, MJW_ERR_HIP, "batch record");
    }
  }
  for (int par = 0; par < 2; ++par)
    if (hipMemcpyAsync(b->d_tail[par], ta[par].data(), ta[par].size() * sizeof(TailArgs),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess)
      BFAIL(b, MJW_ERR_HIP, "tail args");
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, b->stream);
  int done = 0, cur = parity;
  if (use_graph && nsteps >= kGraphSteps) {
    if (cur != 0) {
      if (batch_enqueue_step(b, cur)) BFAIL(b, MJW_ERR_HIP, "batch step");
      cur ^= 1;
      ++done;
    }
    if (nsteps - done >= kGraphSteps) {
      if (!b->gexec) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
          BFAIL(b, MJW_ERR_HIP, "capture");
        int rc = 0;
        for (int s = 0; s < kGraphSteps && !rc; ++s) rc = batch_enqueue_step(b, s & 1);
        hipError_t ee = hipStreamEndCapture(b->stream, &g);
        if (rc || ee != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          BFAIL(b, MJW_ERR_HIP, "batch graph capture");
        }
        ee = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ee != hipSuccess) { b->gexec = nullptr; BFAIL(b, MJW_ERR_HIP, "batch graph"); }
      }
      while (nsteps - done >= kGraphSteps) {
        if (hipGraphLaunch(b->gexec, b->stream) != hipSuccess) BFAIL(b, MJW_ERR_HIP, "graph launch");
        done += kGraphSteps;
      }
    }
  }
  for (; done < nsteps; ++done) {
    if (batch_enqueue_step(b, cur)) BFAIL(b, MJW_ERR_HIP, "batch step");
    cur ^= 1;
  }
  (void)hipEventRecord(e1, b->stream);
  if (hipEventSynchronize(e1) != hipSuccess) BFAIL(b, MJW_ERR_HIP, "batch run failed");
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->ms_step = ms / nsteps;
  for (mjw_problem *p : b->probs) {
    p->cur = cur;
    p->step += nsteps;
  }
  b->last_nsteps = nsteps;
  return MJW_OK;
}

extern "C" int mjw_batch_traces(mjw_batch *b, int iprob, double *traces) {
  if (!b || iprob < 0 || iprob >= (int)b->probs.size() || !traces)
    BFAIL(b, MJW_ERR_ARG, "mjw_batch_traces: bad args");
  mjw_problem *p = b->probs[iprob];
  if (!p->nrec) return MJW_OK;
  if (hipMemcpy(traces, p->d_traces, (size_t)b->last_nsteps * p->nrec * 8,
                hipMemcpyDeviceToHost) != hipSuccess)
    BFAIL(b, MJW_ERR_HIP, "copy");
  return MJW_OK;
}

extern "C" int mjw_batch_timing(const mjw_batch *b, double *ms) {
  if (!b || !ms) BFAIL((mjw_batch *)nullptr, MJW_ERR_ARG, "NULL");
  *ms = b->ms_step;
  return MJW_OK;
}

extern "C" const char *mjw_batch_last_error(const mjw_batch *b) {
  return b ? b->err.c_str() : g_err.c_str();
}

extern "C" int mjw_batch_destroy(mjw_batch *b) {
  if (!b) return MJW_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  batch_free_items(b);
  dfree(b->d_probs);
  dfree(b->d_tail[0]);
  dfree(b->d_tail[1]);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MJW_OK;
}

extern "C" int mjw_tile_choices(const mjw_problem *p, int *out, int cap) {
  if (!p || (!out && cap > 0)) FAIL((mjw_problem *)nullptr, MJW_ERR_ARG, "NULL");
  std::vector<int> v = problem_tiles(p);
  for (int i = 0; i < (int)v.size() && i < cap; ++i) out[i] = v[i];
  return (int)v.size();
}
