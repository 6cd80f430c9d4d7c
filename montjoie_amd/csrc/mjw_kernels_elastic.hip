// Sum-factorised stiffness kernels for isotropic elastodynamics on
// Gauss-Lobatto Q_r hexahedra (and quadrilaterals, k_elastic2d), gfx950.
//
// Restates MltAdd_ElementH1 (Computation/ProdMatVectH1.cxx:674-944) for the
// continuous second-order case: three reference gradients (Rh^T, :745-751),
// physical gradient DF^{-T} grad (:762-777), isotropic stress
// sigma = lambda tr(e) I + 2 mu e (ApplyTensorStiffness,
// Elliptic/Elastic/VarElastic.cxx:22-30), J w DF^{-1} sigma (:797-820), three
// divergences (:839-856), scatter into the component-blocked vector
// (Mesh/NumberMap.cxx:693).
//
// The per-node geometry is stored as M = sqrt(J w) DF^{-1} (9 values, plane-
// major [k][9][E][n*n]): grad' = M^T g = sqrt(Jw) grad u, sigma' = C:grad'
// = sqrt(Jw) sigma, v = M sigma' = J w DF^{-1} sigma -- one 3x3 tensor per
// node instead of J w and J DF^{-1}.  Mapping, colouring, fused leapfrog and
// shot loop as in the scalar kernels (mjw_kernels_scalar.hip).  TRI: for
// straight-sided hexahedra M is rebuilt per node from the element's
// trilinear map (DF^{-1} = adj(DF)/J, so M = sqrt(w/J) adj(DF)) instead of
// read from G.
#include <hip/hip_runtime.h>

#include "mjw_device.hpp"
#include "mjw_internal.hpp"
#include "mjw_tri.hpp"

namespace mjw {

template <int N, int EPB, typename T, typename TG, int MODE, bool TRI>
__global__ __launch_bounds__(EPB *N *N) void k_elastic3d(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    const double *__restrict__ lam_e, const double *__restrict__ mu_e, int64_t E, int64_t e0,
    int64_t e1, const T *__restrict__ U, T *__restrict__ Y, const T *__restrict__ sdiag,
    T alpha, T beta, int64_t N_nodes, int nshots, int nxcd, const TG *__restrict__ coef,
    const TG *__restrict__ glxw) {
  constexpr int P = N * N, NN = N * N * N;
  using A = TG;
  __shared__ T lds[9 * EPB * NN];
  __shared__ TG lc[TRI ? EPB * kCoef : 1];
  const int t = threadIdx.x;
  const int el = t / P, ij = t - el * P;
  const int i = ij % N, j = ij / N;
  const int64_t e = e0 + (int64_t)xcd_block(blockIdx.x, gridDim.x, nxcd) * EPB + el;
  const bool act = e < e1;
  const int64_t eb = e0 + (int64_t)xcd_block(blockIdx.x, gridDim.x, nxcd) * EPB;
  const int64_t plane = E * P;
  const uint32_t lo = (uint32_t)(el * P + ij);
  const uint32_t g_ps = (uint32_t)(plane * sizeof(TG));
  int32_t gid[N];
  A lam = 0, mu = 0;
  TriCol<A> tc;
  if constexpr (TRI) {
    for (int q = t; q < EPB * kCoef; q += blockDim.x)
      lc[q] = eb + q / kCoef < e1 ? coef[eb * kCoef + q] : TG(0);
    __syncthreads();
    if (act) {
      tri_setup<A, TG>(lc + el * kCoef, (A)glxw[i], (A)glxw[j],
                       (A)glxw[N + i] * (A)glxw[N + j], false, tc);
    }
  }
  if (act) {
#pragma unroll
    for (int k = 0; k < N; ++k)
      gid[k] = bld<int32_t>(mkres(map + eb * P), lo * 4u, (uint32_t)(k * plane * 4));
    lam = (A)lam_e[e];
    mu = (A)mu_e[e];
  }
  // LDS planes [9][EPB][NN]: u_m in planes 0..2, the xi_1 fluxes v_{m,0} in
  // planes 3..5, the xi_2 fluxes v_{m,1} in planes 6..8 (written as they are
  // computed, so two barriers per shot); the xi_3 fluxes stay in registers
  T *pl[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) pl[c] = lds + (c * EPB + el) * NN;
  const rsrc_t rs = mkres(sdiag);
  const uint32_t comp_ps = (uint32_t)(N_nodes * sizeof(T));
  // the scatter's previous values and dt^2/M are loaded before the barrier
  // that precedes the divergence, so their latency hides behind it (10% on
  // the Q4 32^3 pass in FP32 and FP64, despite the extra registers)
  constexpr bool PRE = true;
  for (int s = 0; s < nshots; ++s) {
    const rsrc_t ru = mkres(U + (int64_t)s * 3 * N_nodes);
    const rsrc_t ry = mkres(Y + (int64_t)s * 3 * N_nodes);
    T u[3][N];
    if (act) {
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int k = 0; k < N; ++k) {
          u[m][k] = bld<T>(ru, (uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)sizeof(T), m * comp_ps);
          pl[m][k * P + ij] = u[m][k];
        }
    }
    __syncthreads();
    T v2[3][N];
    if (act) {
      T Di[N], Dj[N];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        Di[l] = Dm[i * N + l];
        Dj[l] = Dm[j * N + l];
      }
#pragma unroll
      for (int k = 0; k < N; ++k) {
        A g[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
          for (int l = 0; l < N; ++l) {
            a0 += Di[l] * pl[m][k * P + j * N + l];
            a1 += Dj[l] * pl[m][k * P + l * N + i];
            a2 += Dm[k * N + l] * u[m][l];
          }
          g[m][0] = a0;
          g[m][1] = a1;
          g[m][2] = a2;
        }
        A M[9];
        if constexpr (TRI) {
          tri_m<A>(tc, (A)glxw[k], (A)glxw[N + k], M);
        } else {
          const rsrc_t rg = mkres(G + (k * 9) * plane + eb * P);
#pragma unroll
          for (int c = 0; c < 9; ++c) M[c] = (A)bld<TG>(rg, lo * (uint32_t)sizeof(TG), c * g_ps);
        }
        A du[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int a = 0; a < 3; ++a)
            du[m][a] = M[0 * 3 + a] * g[m][0] + M[1 * 3 + a] * g[m][1] + M[2 * 3 + a] * g[m][2];
        const A tr = du[0][0] + du[1][1] + du[2][2];
        A sg[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int a = 0; a < 3; ++a)
            sg[m][a] = mu * (du[m][a] + du[a][m]) + (m == a ? lam * tr : A(0));
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          pl[3 + m][k * P + ij] = (T)(M[0] * sg[m][0] + M[1] * sg[m][1] + M[2] * sg[m][2]);
          pl[6 + m][k * P + ij] = (T)(M[3] * sg[m][0] + M[4] * sg[m][1] + M[5] * sg[m][2]);
          v2[m][k] = (T)(M[6] * sg[m][0] + M[7] * sg[m][1] + M[8] * sg[m][2]);
        }
      }
    }
    T yk[PRE ? 3 : 1][PRE ? N : 1], sk[PRE && MODE == kLeap ? N : 1];
    if constexpr (PRE) {
      if (act) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const uint32_t off = (uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)sizeof(T);
          if constexpr (MODE == kLeap) sk[k] = bld<T>(rs, off, 0u);
#pragma unroll
          for (int m = 0; m < 3; ++m) yk[m][k] = bld<T>(ry, off, m * comp_ps);
        }
      }
    }
    __syncthreads();
    if (act) {
      T Dti[N], Dtj[N];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        Dti[l] = Dm[l * N + i];
        Dtj[l] = Dm[l * N + j];
      }
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const int32_t mm = gid[k];
        const uint32_t off = (uint32_t)(mm & 0x7fffffff) * (uint32_t)sizeof(T);
        const bool first = mm < 0;
        T sgd = 0;
        if constexpr (MODE == kLeap) sgd = PRE ? sk[PRE ? k : 0] : bld<T>(rs, off, 0u);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          T o = 0;
#pragma unroll
          for (int l = 0; l < N; ++l) {
            o += Dti[l] * pl[3 + m][k * P + j * N + l];
            o += Dtj[l] * pl[6 + m][k * P + l * N + i];
            o += Dm[l * N + k] * v2[m][l];
          }
          const T y = PRE ? yk[PRE ? m : 0][PRE ? k : 0] : bld<T>(ry, off, m * comp_ps);
          T r;
          if (MODE == kLeap) {
            r = first ? (T(2) * u[m][k] - y - sgd * o) : (y - sgd * o);
          } else {
            r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
          }
          bst<T>(r, ry, off, m * comp_ps);
        }
      }
    }
    // the next shot writes planes 0..2 (last read before the barrier above)
    // and planes 3..8 only after its first barrier: no barrier needed here
  }
}

template <int N, int EPB, typename T, typename TG, int MODE, bool TRI>
static void launch_el(const StiffArgs &a, hipStream_t st, unsigned blocks) {
  hipLaunchKernelGGL((k_elastic3d<N, EPB, T, TG, MODE, TRI>), dim3(blocks), dim3(EPB * N * N), 0,
                     st, a.map, static_cast<const TG *>(a.G), static_cast<const T *>(a.Dm), a.lam,
                     a.mu, a.E, a.e0, a.e1, static_cast<const T *>(a.U), static_cast<T *>(a.Y),
                     static_cast<const T *>(a.sdiag), (T)a.alpha, (T)a.beta, a.N, a.nshots,
                     a.nxcd, static_cast<const TG *>(a.coef), static_cast<const TG *>(a.glxw));
}

template <int N, int EPB, typename T, typename TG>
static int run_elastic(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  const unsigned blocks = (unsigned)((ne + EPB - 1) / EPB);
  const bool tri = a.geo == kGeoTrilinear;
  if (tri && (!a.coef || !a.glxw)) return -1002;
  if (a.mode == kLeap) {
    if (tri) launch_el<N, EPB, T, TG, kLeap, true>(a, st, blocks);
    else launch_el<N, EPB, T, TG, kLeap, false>(a, st, blocks);
  } else {
    if (tri) launch_el<N, EPB, T, TG, kApply, true>(a, st, blocks);
    else launch_el<N, EPB, T, TG, kApply, false>(a, st, blocks);
  }
  return (int)hipGetLastError();
}

template <int N, int EPB>
static int run_elastic_prec(const StiffArgs &a, hipStream_t st) {
  if (a.state == kF64) return run_elastic<N, EPB, double, double>(a, st);
  if (a.geom == kG64) return run_elastic<N, EPB, float, double>(a, st);
  return run_elastic<N, EPB, float, float>(a, st);
}

// LDS = 9*EPB*n^3*sizeof(T) (+ coefficients) <= 64 KiB in FP64
#define MJW_TILES_EL(X) \
  X(2, 16) X(2, 32) X(3, 7) X(3, 14) X(4, 4) X(4, 8) X(5, 5) X(5, 3) X(6, 2) X(6, 3) X(7, 1) \
  X(7, 2) X(8, 1)

// ------------------------------------------------------------------ 2-D
// Quadrilaterals, two components: the 2-D instance of the same product
// (ProdMatVectH1.cxx:674-944 on QUADRANGLE_LOBATTO, plane strain
// sigma = lambda tr(e) I + 2 mu e).  One thread per (element, node), EPB
// elements per block; M = sqrt(J w) DF^{-1} (4 values, planes [4][E][n*n]);
// the two state planes of a block are staged in LDS, then overwritten by the
// four flux planes once every thread has read them.  Shots are consecutive
// vectors (u[(s*2 + m)*N + g]) and loop inside the block.
template <int N, int EPB, typename T, typename TG, int MODE>
__global__ __launch_bounds__(EPB *N *N) void k_elastic2d(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    const double *__restrict__ lam_e, const double *__restrict__ mu_e, int64_t E, int64_t e0,
    int64_t e1, const T *__restrict__ U, T *__restrict__ Y, const T *__restrict__ sdiag,
    T alpha, T beta, int64_t N_nodes, int nshots, int nxcd) {
  constexpr int P = N * N, PL = EPB * P;
  using A = TG;
  __shared__ T lds[4 * PL];  // u_0 u_1, then the fluxes f_m,b at plane 2m + b
  const int t = threadIdx.x;
  const int el = t / P, q = t - el * P;
  const int i = q % N, j = q / N;
  const int64_t eb = e0 + (int64_t)xcd_block(blockIdx.x, gridDim.x, nxcd) * EPB;
  const int64_t e = eb + el;
  const bool act = e < e1;
  const int64_t plane = E * P;
  A M[4] = {0, 0, 0, 0};
  A lam = 0, mu = 0;
  int32_t gid = 0;
  if (act) {
    gid = map[e * P + q];
#pragma unroll
    for (int c = 0; c < 4; ++c) M[c] = (A)G[c * plane + e * P + q];
    lam = (A)lam_e[e];
    mu = (A)mu_e[e];
  }
  T Di[N], Dj[N], Dti[N], Dtj[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    Di[l] = Dm[i * N + l];
    Dj[l] = Dm[j * N + l];
    Dti[l] = Dm[l * N + i];
    Dtj[l] = Dm[l * N + j];
  }
  const int64_t g = (int64_t)(gid & 0x7fffffff);
  const bool first = gid < 0;
  const int base = el * P;
  for (int s = 0; s < nshots; ++s) {
    const int64_t o0 = (int64_t)s * 2 * N_nodes + g, o1 = o0 + N_nodes;
    T u0 = 0, u1 = 0, y0 = 0, y1 = 0, sg = 0;
    if (act) {
      u0 = U[o0];
      u1 = U[o1];
      y0 = Y[o0];
      y1 = Y[o1];
      if (MODE == kLeap) sg = sdiag[g];
      lds[base + q] = u0;
      lds[PL + base + q] = u1;
    }
    __syncthreads();
    A f[2][2] = {{0, 0}, {0, 0}};
    if (act) {
      T gr[2][2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        T a0 = 0, a1 = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          a0 += Di[l] * lds[m * PL + base + j * N + l];
          a1 += Dj[l] * lds[m * PL + base + l * N + i];
        }
        gr[m][0] = a0;
        gr[m][1] = a1;
      }
      // du = M^T g (= sqrt(Jw) grad u), sigma' = C : du, f = M sigma'
      A du[2][2];
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int a = 0; a < 2; ++a) du[m][a] = M[0 * 2 + a] * (A)gr[m][0] + M[1 * 2 + a] * (A)gr[m][1];
      const A tr = du[0][0] + du[1][1];
      A sgm[2][2];
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int a = 0; a < 2; ++a) sgm[m][a] = mu * (du[m][a] + du[a][m]) + (m == a ? lam * tr : A(0));
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int b = 0; b < 2; ++b) f[m][b] = M[b * 2 + 0] * sgm[m][0] + M[b * 2 + 1] * sgm[m][1];
    }
    __syncthreads();  // every read of the staged state is done
    if (act) {
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int b = 0; b < 2; ++b) lds[(2 * m + b) * PL + base + q] = (T)f[m][b];
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        T o = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          o += Dti[l] * lds[(2 * m) * PL + base + j * N + l];
          o += Dtj[l] * lds[(2 * m + 1) * PL + base + l * N + i];
        }
        const T u = m ? u1 : u0, y = m ? y1 : y0;
        T r;
        if constexpr (MODE == kLeap)
          r = first ? (T(2) * u - y - sg * o) : (y - sg * o);
        else
          r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
        Y[m ? o1 : o0] = r;
      }
    }
    __syncthreads();  // the fluxes are read before the next shot is staged
  }
}

template <int N, int EPB, typename T, typename TG>
static int run_elastic2d(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  const unsigned blocks = (unsigned)((ne + EPB - 1) / EPB);
#define MJW_EL2(MODE)                                                                           \
  hipLaunchKernelGGL((k_elastic2d<N, EPB, T, TG, MODE>), dim3(blocks), dim3(EPB * N * N), 0, st, \
                     a.map, static_cast<const TG *>(a.G), static_cast<const T *>(a.Dm), a.lam,  \
                     a.mu, a.E, a.e0, a.e1, static_cast<const T *>(a.U), static_cast<T *>(a.Y), \
                     static_cast<const T *>(a.sdiag), (T)a.alpha, (T)a.beta, a.N, a.nshots,      \
                     a.nxcd)
  if (a.mode == kLeap) MJW_EL2(kLeap);
  else MJW_EL2(kApply);
#undef MJW_EL2
  return (int)hipGetLastError();
}

template <int N, int EPB>
static int run_elastic2d_prec(const StiffArgs &a, hipStream_t st) {
  if (a.state == kF64) return run_elastic2d<N, EPB, double, double>(a, st);
  if (a.geom == kG64) return run_elastic2d<N, EPB, float, double>(a, st);
  return run_elastic2d<N, EPB, float, float>(a, st);
}

// about 256 and 512 threads per block
#define MJW_TILES_EL2(X)                                                                       \
  X(2, 64) X(2, 128) X(3, 28) X(3, 56) X(4, 16) X(4, 32) X(5, 10) X(5, 20) X(6, 7) X(6, 14)    \
  X(7, 5) X(7, 10) X(8, 4) X(8, 8)

int launch_stiffness_elastic(const StiffArgs &a, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = a.order + 1;
  if (a.dim == 2) {
#define X(NN, EE) \
  if (n == NN && a.epb == EE) return run_elastic2d_prec<NN, EE>(a, st);
    MJW_TILES_EL2(X)
#undef X
    return -1000;
  }
#define X(NN, EE) \
  if (n == NN && a.epb == EE) return run_elastic_prec<NN, EE>(a, st);
  MJW_TILES_EL(X)
#undef X
  return -1000;
}

std::vector<int> tile_choices_elastic(int dim, int order) {
  std::vector<int> v;
  const int n = order + 1;
#define X(NN, EE) \
  if (n == NN) v.push_back(EE);
  if (dim == 2) {
    MJW_TILES_EL2(X)
  } else {
    MJW_TILES_EL(X)
  }
#undef X
  return v;
}

}  // namespace mjw
