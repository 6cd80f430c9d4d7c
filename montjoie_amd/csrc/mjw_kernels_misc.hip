// Setup and small per-step kernels: FP64 geometry precompute, lumped-mass
// assembly, source injection + receiver interpolation ("tail" of a leapfrog
// step), vector helpers and deterministic FP64 reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "mjw_internal.hpp"

namespace mjw {

int launch_stiffness_scalar(const StiffArgs &a, void *stream);
int launch_stiffness_elastic(const StiffArgs &a, void *stream);
std::vector<int> tile_choices_scalar(int dim, int order);
std::vector<int> tile_choices_elastic(int dim, int order);

int launch_stiffness(const StiffArgs &a, void *stream) {
  if (a.ilv && a.dim == 3 && !a.grad)
    return a.ncomp == 3 ? launch_stiffness_shots_elastic(a, stream)
                        : launch_stiffness_shots(a, stream);
  if (a.ncomp > 1) return launch_stiffness_elastic(a, stream);
  return launch_stiffness_scalar(a, stream);
}

std::vector<int> tile_choices(int dim, int order, int ncomp) {
  if (ncomp > 1) return tile_choices_elastic(dim, order);
  return tile_choices_scalar(dim, order);
}

// first listed column tile; for the 3-D scalar kernels its persistent
// variant when one exists (fastest on the 48^3 Q5 sweep in every precision)
int default_tile(int dim, int order, int ncomp, int state) {
  (void)state;
  std::vector<int> v = tile_choices(dim, order, ncomp);
  if (v.empty()) return 0;
  if (dim == 3 && ncomp == 1)
    for (int t : v)
      if (t == kPipeTile + v[0]) return t;
  return v[0];
}

// ---------------------------------------------------------------------------
// Geometry (FP64).  Per element-local node: DF = d x / d xi from the Q_r
// interpolant of the node positions (exact for multilinear elements; Montjoie
// evaluates DF of the element's geometric transformation, Fb.DFjElem,
// Harmonic/VarGeometryProblem.cxx:693-756), J = det DF, w = tensor GL weight,
// then either the acoustic tensor J w mu DF^{-1} DF^{-T}
// (Elliptic/Helmholtz/VarHelmholtz.cxx:88-117) or the elastic factor
// sqrt(J w) DF^{-1}, and the mass coefficient rho J w (VarHelmholtz.cxx:115).
// ---------------------------------------------------------------------------
__global__ void k_geometry(int dim, int n, int ncomp, int64_t E, const int32_t *__restrict__ map,
                           const double *__restrict__ xyz, const double *__restrict__ D,
                           const double *__restrict__ wgl, const double *__restrict__ rho,
                           const double *__restrict__ mu, const double *__restrict__ rho_q,
                           const double *__restrict__ mu_q, double *__restrict__ G64,
                           float *__restrict__ G32, double *__restrict__ Mloc,
                           int *__restrict__ bad) {
  const int P = n * n;
  const int nd = dim == 2 ? P : P * n;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= E * nd) return;
  const int64_t e = id / nd;
  const int q = (int)(id - e * nd);
  const int i = q % n, j = (q / n) % n, k = q / P;
  auto node = [&](int a, int b, int c) -> int64_t {
    const int ij = a + n * b;
    return (int64_t)(map[((int64_t)c * E + e) * P + ij] & 0x7fffffff);
  };
  double DF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int l = 0; l < n; ++l) {
    const double d0 = D[i * n + l], d1 = D[j * n + l];
    const int64_t g0 = node(l, j, k), g1 = node(i, l, k);
    for (int a = 0; a < dim; ++a) {
      DF[a * dim + 0] += d0 * xyz[g0 * dim + a];
      DF[a * dim + 1] += d1 * xyz[g1 * dim + a];
    }
    if (dim == 3) {
      const double d2 = D[k * n + l];
      const int64_t g2 = node(i, j, l);
      for (int a = 0; a < 3; ++a) DF[a * 3 + 2] += d2 * xyz[g2 * 3 + a];
    }
  }
  double J, Di[9];
  if (dim == 2) {
    J = DF[0] * DF[3] - DF[1] * DF[2];
    Di[0] = DF[3] / J; Di[1] = -DF[1] / J; Di[2] = -DF[2] / J; Di[3] = DF[0] / J;
  } else {
    const double c00 = DF[4] * DF[8] - DF[5] * DF[7];
    const double c01 = DF[5] * DF[6] - DF[3] * DF[8];
    const double c02 = DF[3] * DF[7] - DF[4] * DF[6];
    J = DF[0] * c00 + DF[1] * c01 + DF[2] * c02;
    Di[0] = c00 / J; Di[1] = (DF[2] * DF[7] - DF[1] * DF[8]) / J; Di[2] = (DF[1] * DF[5] - DF[2] * DF[4]) / J;
    Di[3] = c01 / J; Di[4] = (DF[0] * DF[8] - DF[2] * DF[6]) / J; Di[5] = (DF[2] * DF[3] - DF[0] * DF[5]) / J;
    Di[6] = c02 / J; Di[7] = (DF[1] * DF[6] - DF[0] * DF[7]) / J; Di[8] = (DF[0] * DF[4] - DF[1] * DF[3]) / J;
  }
  if (!(J > 0.0)) atomicOr(bad, 1);
  const double w = wgl[i] * wgl[j] * (dim == 3 ? wgl[k] : 1.0);
  const double Jw = J * w;
  // media at this quadrature point: per-node fields when given, else the
  // element's constants
  if (Mloc) Mloc[id] = (rho_q ? rho_q[id] : rho[e]) * Jw;
  const int ij = i + n * j;
  if (ncomp == 1) {
    const double c = Jw * (mu_q ? mu_q[id] : mu[e]);
    if (dim == 2) {
      double g[3];
      g[0] = c * (Di[0] * Di[0] + Di[1] * Di[1]);
      g[1] = c * (Di[0] * Di[2] + Di[1] * Di[3]);
      g[2] = c * (Di[2] * Di[2] + Di[3] * Di[3]);
      for (int cc = 0; cc < 3; ++cc) {
        const int64_t o = ((int64_t)cc * E + e) * P + ij;
        if (G64) G64[o] = g[cc];
        if (G32) G32[o] = (float)g[cc];
      }
    } else {
      double g[6];
      const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
      for (int cc = 0; cc < 6; ++cc) {
        const double *ra = Di + 3 * ia[cc], *rb = Di + 3 * ib[cc];
        g[cc] = c * (ra[0] * rb[0] + ra[1] * rb[1] + ra[2] * rb[2]);
      }
      for (int cc = 0; cc < 6; ++cc) {
        const int64_t o = (((int64_t)k * 6 + cc) * E + e) * P + ij;
        if (G64) G64[o] = g[cc];
        if (G32) G32[o] = (float)g[cc];
      }
    }
  } else {
    // elastic: M = sqrt(J w) DF^{-1}, dim*dim planes per layer k
    const double sq = sqrt(Jw);
    const int dd = dim * dim;
    for (int cc = 0; cc < dd; ++cc) {
      const int64_t o = (((int64_t)k * dd + cc) * E + e) * P + ij;
      const double v = sq * Di[cc];
      if (G64) G64[o] = v;
      if (G32) G32[o] = (float)v;
    }
  }
}

int launch_geometry_full(int dim, int order, int ncomp, int64_t E, const int32_t *map,
                         const double *xyz, const double *D, const double *w, const double *rho,
                         const double *mu, const double *rho_q, const double *mu_q, double *G64,
                         float *G32, double *Mloc, int *bad, void *stream) {
  const int n = order + 1;
  const int nd = dim == 2 ? n * n : n * n * n;
  const int64_t tot = E * nd;
  const int bs = 256;
  hipLaunchKernelGGL(k_geometry, dim3((unsigned)((tot + bs - 1) / bs)), dim3(bs), 0,
                     static_cast<hipStream_t>(stream), dim, n, ncomp, E, map, xyz, D, w, rho, mu,
                     rho_q, mu_q, G64, G32, Mloc, bad);
  return (int)hipGetLastError();
}

// values of the state at probe points: out[(s*npts + r)*ncomp + m] =
// sum_q w[r][q] u_m(s)[gid[r][q]] (FP64 sums in a fixed order); one thread
// per (point, shot, component)
template <typename T>
__global__ void k_probe(const T *__restrict__ U, int ilv, int nshots, int ncomp, int64_t N,
                        int npts, int nd, const int32_t *__restrict__ gid,
                        const double *__restrict__ w, double *__restrict__ out) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)npts * nshots * ncomp;
  if (id >= tot) return;
  const int m = (int)(id % ncomp);
  const int64_t rs = id / ncomp;
  const int r = (int)(rs % npts), s = (int)(rs / npts);
  const int64_t NT = (int64_t)ncomp * N;
  double acc = 0.0;
  for (int q = 0; q < nd; ++q) {
    const double wq = w[(int64_t)r * nd + q];
    if (wq == 0.0) continue;
    const int64_t d = (int64_t)m * N + gid[(int64_t)r * nd + q];
    acc += wq * (double)U[ilv ? d * nshots + s : (int64_t)s * NT + d];
  }
  out[id] = acc;
}

int launch_probe(int state, const void *U, int ilv, int nshots, int ncomp, int64_t N, int npts,
                 int nd, const int32_t *gid, const double *w, double *out, void *stream) {
  const int64_t tot = (int64_t)npts * nshots * ncomp;
  if (tot <= 0) return 0;
  const dim3 grid((unsigned)((tot + 255) / 256)), block(256);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_probe<double>, grid, block, 0, st, static_cast<const double *>(U), ilv,
                       nshots, ncomp, N, npts, nd, gid, w, out);
  else
    hipLaunchKernelGGL(k_probe<float>, grid, block, 0, st, static_cast<const float *>(U), ilv,
                       nshots, ncomp, N, npts, nd, gid, w, out);
  return (int)hipGetLastError();
}

// lumped mass of one colour: first toucher writes, others add (deterministic)
__global__ void k_mass(const int32_t *__restrict__ map, const double *__restrict__ Mloc,
                       int64_t E, int n, int dim, int64_t e0, int64_t e1, double *__restrict__ M) {
  const int P = n * n, nd = dim == 2 ? P : P * n;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (e1 - e0) * nd) return;
  const int64_t e = e0 + id / nd;
  const int q = (int)(id % nd);
  const int ij = q % P, k = q / P;
  const int32_t m = map[((int64_t)k * E + e) * P + ij];
  const int64_t g = m & 0x7fffffff;
  const double v = Mloc[e * nd + q];
  if (m < 0)
    M[g] = v;
  else
    M[g] += v;
}

int launch_mass_assemble(const int32_t *map, const double *Mloc, int64_t E, int dim, int order,
                         int64_t e0, int64_t e1, double *M, void *stream) {
  const int n = order + 1;
  const int nd = dim == 2 ? n * n : n * n * n;
  const int64_t tot = (e1 - e0) * nd;
  if (tot <= 0) return 0;
  hipLaunchKernelGGL(k_mass, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), map, Mloc, E, n, dim, e0, e1, M);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tail of leapfrog step n (one block): Dirac source injection into U^{n+2}
// being formed (AddScalarSourceAtTime at t_n + dt, VarInstationary.cxx:
// 2391-2409, scaled by dt^2/M as the solve does, ModifiedEquation.cxx:316-333)
// and receiver interpolation of U^{n+1}, the iterate the next step hands to
// GiveIterate (VarInstationary.cxx:798-853).  The step index lives on the
// device so a captured graph can be replayed.
// ---------------------------------------------------------------------------
constexpr int kTailThreads = 1024;

// receiver values of the iterate U: one wavefront per (receiver, shot) pair
// it = wave0, wave0 + stride, ..., lanes over the element's nodes, reduced in
// a fixed order (the result does not depend on the launch shape)
template <typename T>
__device__ __forceinline__ void record_waves(const TailArgs &a, int ncomp, int64_t step,
                                             int64_t base, int64_t wave0, int64_t stride) {
  const T *U = static_cast<const T *>(a.U);
  const int lane = threadIdx.x & 63;
  const int64_t NT = (int64_t)ncomp * a.N;
  const int64_t tot = (int64_t)a.nrec * a.nshots;
  for (int64_t it = wave0; it < tot; it += stride) {
    const int r = (int)(it % a.nrec), s = (int)(it / a.nrec);
    const int64_t cbase = (int64_t)a.rec_comp[r] * a.N;
    double acc = 0.0;
    for (int q = lane; q < a.nd; q += 64) {
      const int64_t d = cbase + a.rec_gid[(int64_t)r * a.nd + q];
      acc += a.rec_w[(int64_t)r * a.nd + q] *
             (double)U[a.ilv ? d * a.nshots + s : (int64_t)s * NT + d];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) a.traces[((step - base + a.row_offset) * a.nshots + s) * a.nrec + r] = acc;
  }
}

// many receivers: the recording spread over the chip, before the one-block
// tail (which then only injects the sources and advances the step counter;
// it runs after this kernel on the same stream, so the step read here is
// the one the tail advances)
constexpr int kRecordThreads = 256;
template <typename T>
__global__ __launch_bounds__(kRecordThreads) void k_record(TailArgs a, int ncomp) {
  const int64_t nw = kRecordThreads / 64;
  record_waves<T>(a, ncomp, a.ctl[0], a.ctl[1], (int64_t)blockIdx.x * nw + (threadIdx.x >> 6),
                  (int64_t)gridDim.x * nw);
}

template <typename T>
__device__ __forceinline__ void tail_body(const TailArgs &a, int ncomp) {
  const int64_t step = a.ctl[0];
  const int64_t base = a.ctl[1];
  T *P = static_cast<T *>(a.P);
  const T *sd = static_cast<const T *>(a.sdiag);
  const double pulse = (a.pulse && step < a.npulse) ? a.pulse[step] : 0.0;
  const int64_t NT = (int64_t)ncomp * a.N;
  if (pulse != 0.0 && P) {
    for (int idx = threadIdx.x; idx < a.nsrc; idx += blockDim.x) {
      const int64_t dof = a.src_dof[idx];
      const int s = a.src_shot[idx];
      const int64_t g = dof % a.N;
      T *y = P + (a.ilv ? dof * a.nshots + s : (int64_t)s * NT + dof);
      // sdiag = null: the source enters unscaled (first-order system)
      *y += sd ? sd[g] * (T)(pulse * a.src_val[idx]) : (T)(pulse * a.src_val[idx]);
    }
  }
  if (a.traces) {
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    record_waves<T>(a, ncomp, step, base, wave, nw);
  }
  if (a.nadj && P) {
    // adjoint source: receiver transpose of dJ/dtrace at this step, unique
    // nodes in CSR form so the sum per node has a fixed order
    const double *gr = a.adj_g + (step + a.adj_row) * a.adj_nrec;
    T *Ps = P + (int64_t)a.adj_shot * NT;
    for (int idx = threadIdx.x; idx < a.nadj; idx += blockDim.x) {
      const int32_t node = a.adj_node[idx];
      double v = 0.0;
      for (int k = a.adj_ptr[idx]; k < a.adj_ptr[idx + 1]; ++k) v += a.adj_w[k] * gr[a.adj_rec[k]];
      Ps[node] += sd[node] * (T)v;
    }
  }
  __syncthreads();
  if (a.advance && threadIdx.x == 0) a.ctl[0] = step + a.advance;
}

template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_tail(TailArgs a, int ncomp) {
  tail_body<T>(a, ncomp);
}

template <typename T>
__global__ __launch_bounds__(kTailThreads) void k_tail_batch(const TailArgs *__restrict__ args) {
  tail_body<T>(args[blockIdx.x], 1);
}

int launch_tail_batch(int state, const TailArgs *args, int nprob, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_tail_batch<double>, dim3(nprob), dim3(kTailThreads), 0, st, args);
  else
    hipLaunchKernelGGL(k_tail_batch<float>, dim3(nprob), dim3(kTailThreads), 0, st, args);
  return (int)hipGetLastError();
}

int launch_tail_n(const TailArgs &a, int ncomp, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  // more (receiver, shot) pairs than four per wave of the one-block tail:
  // record them with a grid of their own first
  const int64_t tot = (int64_t)a.nrec * a.nshots;
  TailArgs t = a;
  if (a.traces && tot > 4 * (kTailThreads / 64)) {
    const int64_t per = kRecordThreads / 64;
    const unsigned grid = (unsigned)std::min<int64_t>((tot + per - 1) / per, 2048);
    if (a.state == kF64)
      hipLaunchKernelGGL(k_record<double>, dim3(grid), dim3(kRecordThreads), 0, st, a, ncomp);
    else
      hipLaunchKernelGGL(k_record<float>, dim3(grid), dim3(kRecordThreads), 0, st, a, ncomp);
    t.traces = nullptr;
  }
  if (a.state == kF64)
    hipLaunchKernelGGL(k_tail<double>, dim3(1), dim3(kTailThreads), 0, st, t, ncomp);
  else
    hipLaunchKernelGGL(k_tail<float>, dim3(1), dim3(kTailThreads), 0, st, t, ncomp);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------ vector helpers
template <typename T>
__global__ void k_axpby(int64_t n, double a, const T *__restrict__ x, double b, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (T)(b == 0.0 ? a * (double)x[i] : a * (double)x[i] + b * (double)y[i]);
}

static unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

int launch_axpby(int state, int64_t n, double a, const void *x, double b, void *y, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid_for(n)), dim3(256), 0, st, n, a,
                       (const double *)x, b, (double *)y);
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid_for(n)), dim3(256), 0, st, n, a,
                       (const float *)x, b, (float *)y);
  return (int)hipGetLastError();
}

template <typename A, typename B>
__global__ void k_convert(int64_t n, const A *__restrict__ x, B *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (B)x[i];
}

int launch_convert_to_f64(int state, int64_t n, const void *x, double *y, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    return (int)hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, st);
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(256), 0, st, n,
                     (const float *)x, y);
  return (int)hipGetLastError();
}

int launch_convert_from_f64(int state, int64_t n, const double *x, void *y, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    return (int)hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, st);
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, st, n, x,
                     (float *)y);
  return (int)hipGetLastError();
}

// y[i] *= c*d[node(i)], node(i) = (i / div) % N (div = S interleaved, else 1)
template <typename T>
__global__ void k_scale_by(int64_t n, int64_t N, int64_t div, const double *__restrict__ d,
                           double c, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (T)((double)y[i] * (c * d[(i / div) % N]));
}

int launch_scale_by(int state, int64_t n, int S, int ilv, int64_t N, const double *d, double c,
                    void *y, void *stream) {
  const int64_t div = ilv ? S : 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_scale_by<double>, dim3(grid_for(n)), dim3(256), 0, st, n, N, div, d, c,
                       (double *)y);
  else
    hipLaunchKernelGGL(k_scale_by<float>, dim3(grid_for(n)), dim3(256), 0, st, n, N, div, d, c,
                       (float *)y);
  return (int)hipGetLastError();
}

template <typename T>
__global__ void k_mask(int64_t n, int64_t N, int64_t div, const uint8_t *__restrict__ dir,
                       T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (dir[(i / div) % N]) y[i] = T(0);
}

int launch_mask(int state, int64_t n, int S, int ilv, int64_t N, const uint8_t *dir, void *y,
                void *stream) {
  if (!dir) return 0;
  const int64_t div = ilv ? S : 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_mask<double>, dim3(grid_for(n)), dim3(256), 0, st, n, N, div, dir,
                       (double *)y);
  else
    hipLaunchKernelGGL(k_mask<float>, dim3(grid_for(n)), dim3(256), 0, st, n, N, div, dir,
                       (float *)y);
  return (int)hipGetLastError();
}

// Node-wise steps of the fourth-order modified-equation leapfrog
// (ModifiedEquationIterator::Advance with order 4 and SetInitialConditionS,
// Instationary/ModifiedEquation.cxx:124-170,347-407), on whole state vectors
// with the per-node mass D, damping S (null: none) and Dirichlet flags; all
// arithmetic in FP64.  Operands by op (y is written):
//   kMeVel    y = c0 (3 a - 4 b + x)                      (dt^2 dU/dt, c0 = dt/2)
//   kMeSolve  y = mask(D^{-1} (a - S b))                  (b null: no S term)
//   kMeSub    y = a - S b
//   kMeNew    y = mask((D + c0 S)^{-1} (a + c1 b + c2 S x + 2 D u - (D - c0 S) w))
//   kMePre    y = a - 2 b + x
//   kMeTaylor y = w + c0 (c1 a + c0 (b/2 + c0 (c1 x/6 + c0 u/24))), c1 = +-1,
//             zero on the Dirichlet nodes when c2 != 0 (SetInitialConditionS
//             masks U^n and U^{n+1} but leaves U^{n-1} as it is)
template <typename T>
__global__ void k_me4(int op, int64_t n, int64_t N, int64_t div, const double *__restrict__ D,
                      const double *__restrict__ Sd, const uint8_t *__restrict__ dir,
                      const T *__restrict__ a, const T *__restrict__ b, const T *__restrict__ x,
                      const T *__restrict__ u, const T *__restrict__ w, double c0, double c1,
                      double c2, T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = (i / div) % N;
    const double s = Sd ? Sd[g] : 0.0;
    double r;
    switch (op) {
      case kMeVel:
        r = c0 * (3.0 * (double)a[i] - 4.0 * (double)b[i] + (double)x[i]);
        break;
      case kMeSolve:
        r = dir[g] ? 0.0 : ((double)a[i] - (b ? s * (double)b[i] : 0.0)) / D[g];
        break;
      case kMeSub:
        r = (double)a[i] - s * (double)b[i];
        break;
      case kMePre:
        r = (double)a[i] - 2.0 * (double)b[i] + (double)x[i];
        break;
      case kMeNew: {
        const double dg = D[g];
        const double rhs = (double)a[i] + c1 * (double)b[i] + (x ? c2 * s * (double)x[i] : 0.0) +
                           2.0 * dg * (double)u[i] - (dg - c0 * s) * (double)w[i];
        r = dir[g] ? 0.0 : rhs / (dg + c0 * s);
        break;
      }
      default: {  // kMeTaylor
        const double t4 = c0 * (double)u[i] / 24.0;
        const double t3 = c0 * (c1 * (double)x[i] / 6.0 + t4);
        const double t2 = c0 * ((double)b[i] / 2.0 + t3);
        r = (c2 != 0.0 && dir[g]) ? 0.0 : (double)w[i] + c0 * (c1 * (double)a[i] + t2);
        break;
      }
    }
    y[i] = (T)r;
  }
}

int launch_me4(int state, int op, int64_t n, int S, int ilv, int64_t N, const double *D,
               const double *Sd, const uint8_t *dir, const void *a, const void *b, const void *x,
               const void *u, const void *w, double c0, double c1, double c2, void *y,
               void *stream) {
  const int64_t div = ilv ? S : 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_me4<double>, dim3(grid_for(n)), dim3(256), 0, st, op, n, N, div, D, Sd,
                       dir, (const double *)a, (const double *)b, (const double *)x,
                       (const double *)u, (const double *)w, c0, c1, c2, (double *)y);
  else
    hipLaunchKernelGGL(k_me4<float>, dim3(grid_for(n)), dim3(256), 0, st, op, n, N, div, D, Sd,
                       dir, (const float *)a, (const float *)b, (const float *)x,
                       (const float *)u, (const float *)w, c0, c1, c2, (float *)y);
  return (int)hipGetLastError();
}

// caller layout [S][ncomp][N] <-> solver layout (see mjw_internal.hpp)
__device__ __forceinline__ int64_t solver_index(int64_t s, int64_t mN_g, int S, int64_t per,
                                                int ilv) {
  return ilv ? mN_g * S + s : s * per + mN_g;
}

template <typename T>
__global__ void k_to_solver(int S, int ncomp, int64_t N, int ilv, const int32_t *__restrict__ ninv,
                            const T *__restrict__ x, T *__restrict__ y) {
  const int64_t per = (int64_t)ncomp * N, n = per * S;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
       u += (int64_t)gridDim.x * blockDim.x) {
    // u runs over the caller's layout; the value of solver node i = nperm[g]
    // comes from caller node ninv[i]: write solver slot of (s, m, i)
    const int64_t s = u / per, r = u - s * per, m = r / N, i = r - m * N;
    y[solver_index(s, m * N + i, S, per, ilv)] = x[s * per + m * N + ninv[i]];
  }
}

template <typename T>
__global__ void k_from_solver(int S, int ncomp, int64_t N, int ilv,
                              const int32_t *__restrict__ nperm, const T *__restrict__ x,
                              T *__restrict__ y) {
  const int64_t per = (int64_t)ncomp * N, n = per * S;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n;
       u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = u / per, r = u - s * per, m = r / N, g = r - m * N;
    y[u] = x[solver_index(s, m * N + nperm[g], S, per, ilv)];
  }
}

// One caller vector (FP64, caller's node numbering, [ncomp][N]) into shot s
// of the solver layout in the state precision, and back: the renumbering and
// the precision change of the host-vector entry points, on the device.
template <typename TO>
__global__ void k_to_solver_f64(int S, int s, int ncomp, int64_t N, int ilv,
                                const int32_t *__restrict__ ninv, const double *__restrict__ x,
                                TO *__restrict__ y) {
  const int64_t per = (int64_t)ncomp * N;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < per;
       u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = u / N, i = u - m * N;
    y[solver_index(s, m * N + i, S, per, ilv)] = (TO)x[m * N + ninv[i]];
  }
}

template <typename TI>
__global__ void k_from_solver_f64(int S, int s, int ncomp, int64_t N, int ilv,
                                  const int32_t *__restrict__ nperm, const TI *__restrict__ x,
                                  double *__restrict__ y) {
  const int64_t per = (int64_t)ncomp * N;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < per;
       u += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = u / N, g = u - m * N;
    y[u] = (double)x[solver_index(s, m * N + nperm[g], S, per, ilv)];
  }
}

int launch_to_solver_f64(int state, int S, int s, int ncomp, int64_t N, int ilv,
                         const int32_t *ninv, const double *x, void *y, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = (int64_t)ncomp * N;
  if (state == kF64)
    hipLaunchKernelGGL(k_to_solver_f64<double>, dim3(grid_for(n)), dim3(256), 0, st, S, s, ncomp,
                       N, ilv, ninv, x, (double *)y);
  else
    hipLaunchKernelGGL(k_to_solver_f64<float>, dim3(grid_for(n)), dim3(256), 0, st, S, s, ncomp,
                       N, ilv, ninv, x, (float *)y);
  return (int)hipGetLastError();
}

int launch_from_solver_f64(int state, int S, int s, int ncomp, int64_t N, int ilv,
                           const int32_t *nperm, const void *x, double *y, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = (int64_t)ncomp * N;
  if (state == kF64)
    hipLaunchKernelGGL(k_from_solver_f64<double>, dim3(grid_for(n)), dim3(256), 0, st, S, s,
                       ncomp, N, ilv, nperm, (const double *)x, y);
  else
    hipLaunchKernelGGL(k_from_solver_f64<float>, dim3(grid_for(n)), dim3(256), 0, st, S, s,
                       ncomp, N, ilv, nperm, (const float *)x, y);
  return (int)hipGetLastError();
}

int launch_to_solver(int state, int S, int ncomp, int64_t N, int ilv, const int32_t *ninv,
                     const void *x, void *y, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = (int64_t)S * ncomp * N;
  if (state == kF64)
    hipLaunchKernelGGL(k_to_solver<double>, dim3(grid_for(n)), dim3(256), 0, st, S, ncomp, N, ilv,
                       ninv, (const double *)x, (double *)y);
  else
    hipLaunchKernelGGL(k_to_solver<float>, dim3(grid_for(n)), dim3(256), 0, st, S, ncomp, N, ilv,
                       ninv, (const float *)x, (float *)y);
  return (int)hipGetLastError();
}

int launch_from_solver(int state, int S, int ncomp, int64_t N, int ilv, const int32_t *nperm,
                       const void *x, void *y, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t n = (int64_t)S * ncomp * N;
  if (state == kF64)
    hipLaunchKernelGGL(k_from_solver<double>, dim3(grid_for(n)), dim3(256), 0, st, S, ncomp, N,
                       ilv, nperm, (const double *)x, (double *)y);
  else
    hipLaunchKernelGGL(k_from_solver<float>, dim3(grid_for(n)), dim3(256), 0, st, S, ncomp, N,
                       ilv, nperm, (const float *)x, (float *)y);
  return (int)hipGetLastError();
}

// s = dt^2 / M on free nodes, 0 on Dirichlet nodes
template <typename T>
__global__ void k_sdiag(int64_t N, const double *__restrict__ M, const uint8_t *__restrict__ dir,
                        double dt2, T *__restrict__ s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    s[i] = (dir && dir[i]) ? T(0) : (T)(dt2 / M[i]);
}

// step counter and trace base row without a host copy: two lanes, vector stores
__global__ void k_set_ctl(int64_t *__restrict__ ctl, int64_t step, int64_t base) {
  const int i = threadIdx.x;
  if (i < 2) ctl[i] = i == 0 ? step : base;
}

int launch_set_ctl(int64_t *ctl, int64_t step, int64_t base, void *stream) {
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), ctl,
                     step, base);
  return (int)hipGetLastError();
}

int launch_make_sdiag(int state, int64_t N, const double *M, const uint8_t *dir, double dt2,
                      void *s, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_sdiag<double>, dim3(grid_for(N)), dim3(256), 0, st, N, M, dir, dt2,
                       (double *)s);
  else
    hipLaunchKernelGGL(k_sdiag<float>, dim3(grid_for(N)), dim3(256), 0, st, N, M, dir, dt2,
                       (float *)s);
  return (int)hipGetLastError();
}

// Damped leapfrog (S_h != 0, Advance, ModifiedEquation.cxx:295-345):
//   (D + dt/2 S) U^{n+1} = 2 D U^n - (D - dt/2 S) U^{n-1} + dt^2 (F - A U^n).
// The colour passes form Y* = 2 U^n - U^{n-1} + s (F - A U^n) with
// s = dt^2 / (D + dt/2 S); the exact update is Y* + k (U^{n-1} - U^n) with
// k = dt S / (D + dt/2 S), nonzero only on damped nodes.  Over the damped
// nodes idx (every component and shot):
//   mode 0: save = Y (U^{n-1}, before the passes overwrite it)
//   mode 1: Y += k (save - U)
//   mode 2: Y += coef * U
template <typename T>
__global__ void k_damp(int mode, int64_t ndamp, int ncomp, int S, int ilv, int64_t N,
                       const int32_t *__restrict__ idx, const double *__restrict__ coef,
                       T *__restrict__ y, T *__restrict__ save, const T *__restrict__ u) {
  const int64_t tot = ndamp * ncomp * S;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < tot;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = j % ndamp, rest = j / ndamp;
    const int64_t m = rest % ncomp, s = rest / ncomp;
    const int64_t g = idx[i];
    const int64_t o = ilv ? ((m * N + g) * S + s) : (s * ncomp * N + m * N + g);
    if (mode == 0)
      save[j] = y[o];
    else if (mode == 1)
      y[o] += (T)coef[i] * (save[j] - u[o]);
    else
      y[o] += (T)coef[i] * u[o];
  }
}

int launch_damp(int state, int mode, int64_t ndamp, int ncomp, int S, int ilv, int64_t N,
                const int32_t *idx, const double *coef, void *y, void *save, const void *u,
                void *stream) {
  const int64_t tot = ndamp * ncomp * S;
  if (tot <= 0) return 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL(k_damp<double>, dim3(grid_for(tot)), dim3(256), 0, st, mode, ndamp, ncomp,
                       S, ilv, N, idx, coef, (double *)y, (double *)save, (const double *)u);
  else
    hipLaunchKernelGGL(k_damp<float>, dim3(grid_for(tot)), dim3(256), 0, st, mode, ndamp, ncomp,
                       S, ilv, N, idx, coef, (float *)y, (float *)save, (const float *)u);
  return (int)hipGetLastError();
}

// G[(plane*E + e)*P + q] *= sc[e] over all planes (media update)
template <typename T>
__global__ void k_scale_planes(T *__restrict__ G, int64_t n, int64_t E, int P,
                               const double *__restrict__ sc) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    G[i] = (T)((double)G[i] * sc[(i / P) % E]);
}

int launch_scale_planes(int g32, void *G, int64_t nplanes, int64_t E, int P, const double *sc,
                        void *stream) {
  const int64_t n = nplanes * E * P;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (g32)
    hipLaunchKernelGGL(k_scale_planes<float>, dim3(grid_for(n)), dim3(256), 0, st, (float *)G, n,
                       E, P, sc);
  else
    hipLaunchKernelGGL(k_scale_planes<double>, dim3(grid_for(n)), dim3(256), 0, st, (double *)G,
                       n, E, P, sc);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------- reductions
// Deterministic two-pass FP64 reduction per shot: fixed grid, per-block
// partials summed by one block in a fixed order.
constexpr int kRedBlocks = 1024;

template <typename T, int DIFF>
__global__ __launch_bounds__(256) void k_dot_partial(int64_t per_shot, int64_t N, int ilv,
                                                     const double *__restrict__ w,
                                                     const T *__restrict__ x,
                                                     const T *__restrict__ y,
                                                     double *__restrict__ part) {
  // shot s, entry i (= m*N + g): x[i*S + s] when interleaved, x[s*per + i] else
  const int s = blockIdx.y, S = gridDim.y;
  const int64_t base = ilv ? s : (int64_t)s * per_shot, stride = ilv ? S : 1;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < per_shot;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double wi = w ? w[i % N] : 1.0;
    const int64_t k = base + i * stride;
    if (DIFF) {
      const double d = (double)x[k] - (double)y[k];
      acc += wi * d * d;
    } else {
      acc += wi * (double)x[k] * (double)y[k];
    }
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)s * gridDim.x + blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_dot_final(int nb, const double *__restrict__ part,
                                                   double *__restrict__ out) {
  const int s = blockIdx.x;
  __shared__ double sh[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) acc += part[(int64_t)s * nb + i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[s] = sh[0];
}

template <int DIFF>
static int dot_impl(int state, int64_t N, int ncomp, int S, int ilv, const double *w,
                    const void *x, const void *y, double *out, double *part, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t per = (int64_t)ncomp * N;
  int nb = (int)std::min<int64_t>(kRedBlocks, (per + 255) / 256);
  if (nb < 1) nb = 1;
  if (state == kF64)
    hipLaunchKernelGGL((k_dot_partial<double, DIFF>), dim3(nb, S), dim3(256), 0, st, per, N, ilv,
                       w, (const double *)x, (const double *)y, part);
  else
    hipLaunchKernelGGL((k_dot_partial<float, DIFF>), dim3(nb, S), dim3(256), 0, st, per, N, ilv,
                       w, (const float *)x, (const float *)y, part);
  hipLaunchKernelGGL(k_dot_final, dim3(S), dim3(256), 0, st, nb, part, out);
  return (int)hipGetLastError();
}

int launch_dot_part(int state, int64_t N, int ncomp, int S, int ilv, const double *w,
                    const void *x, const void *y, double *out, double *part, void *stream) {
  return dot_impl<0>(state, N, ncomp, S, ilv, w, x, y, out, part, stream);
}
int launch_dot_diff_part(int state, int64_t N, int ncomp, int S, int ilv, const double *w,
                         const void *a, const void *b, double *out, double *part, void *stream) {
  return dot_impl<1>(state, N, ncomp, S, ilv, w, a, b, out, part, stream);
}
int reduction_scratch_doubles(int S) { return kRedBlocks * S; }

}  // namespace mjw
