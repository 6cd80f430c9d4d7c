// Sum-factorised stiffness kernels for the scalar (acoustic) wave equation on
// Gauss-Lobatto Q_r quadrilaterals / hexahedra, gfx950.
//
// Restates, per element, MltAdd_ElementH1_Scalar (Computation/
// ProdMatVectScalarH1.cxx:256-668) on the Lobatto path: gather (NumberMap.cxx:
// 661-699), reference gradient Rh^T u (HexahedronLobatto.cxx:257-281), per-node
// tensor J w DF^{-1} mu DF^{-T} (ProdMatVectScalarH1.cxx:472-495), divergence
// Rh v (HexahedronLobatto.cxx:217-251), scatter-add (NumberMap.cxx:745-777).
//
// Layout / mapping (MI355X):
//  * elements are sorted by colour; one launch covers one colour, so the
//    scatter has no conflicts and needs no atomics;
//  * 3-D: one thread per (element, i, j) column, looping k; the column's
//    values stay in registers (xi_3 contractions) and the element is staged
//    in LDS for the xi_1 / xi_2 contractions.  2-D: one thread per node;
//  * map and geometry are plane-major ([k][comp][E][n*n]) so each wavefront
//    load is one contiguous run;
//  * kLeap fuses the lumped-mass leapfrog update: the first element to touch
//    a node (bit 31 of its map entry, lowest colour) writes
//    Y = 2U - Y - s*out, later ones Y -= s*out, with s = dt^2/M (0 on
//    Dirichlet nodes), so U^{n+1} overwrites U^{n-1} in the same pass;
//  * kApply computes Y = beta*Y + alpha*A*U (first toucher applies beta);
//  * several shots (right-hand sides) share the mesh: the geometry of the
//    column is loaded once into registers and reused for every shot (shot-
//    major state, each shot a contiguous vector).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mjw_device.hpp"
#include "mjw_internal.hpp"
#include "mjw_tri.hpp"

namespace mjw {

template <typename T, typename TG>
struct Mix {  // geometry product in the geometry's precision
  using A = TG;
};

// Body of the 3-D scalar stiffness pass for the elements [eb, min(eb+EPB, e1))
// of one colour; threads t >= EPB*n^2 of the block idle (they only meet the
// barriers).  lu, lv: 2*EPB*n^3 values of LDS; gpart: EPB*n^2 doubles (GRAD).
// TRI: the per-node tensor is rebuilt from the element's trilinear map
// (coef: kCoef values per element, staged through lc = EPB*kCoef values of
// LDS; glxw: GL points then weights) instead of read from G.
template <int N, int EPB, typename T, typename TG, int MODE, bool HOIST, bool GRAD, bool TRI>
__device__ __forceinline__ void scalar3d_body(
    T *lu, T *lv, double *gpart, TG *lc, const int32_t *__restrict__ map,
    const TG *__restrict__ G, const T *__restrict__ Dm, int64_t E, int64_t eb, int64_t e1,
    const T *__restrict__ U, T *__restrict__ Y, const T *__restrict__ sdiag, T alpha, T beta,
    int64_t N_nodes, int nshots, double *__restrict__ grad, const double *__restrict__ mu_e,
    const TG *__restrict__ coef, const TG *__restrict__ glxw) {
  // GRAD (adjoint step, nshots == 2: shot 0 = forward iterate U^j rebuilt
  // backwards, shot 1 = adjoint nu_j): grad[e] -= (A_e nu_e . U_e) / mu_e
  constexpr int P = N * N, NN = N * N * N;
  using A = typename Mix<T, TG>::A;
  T ufw[GRAD ? N : 1];  // shot-0 values of the column (forward iterate)
  double gacc = 0.0;
  const int t = threadIdx.x;
  const int el = t / P, ij = t - el * P;
  const int i = ij % N, j = ij / N;
  // block-uniform resources; per-lane byte offsets are 32-bit
  const int64_t plane = E * P;
  const uint32_t lo = (uint32_t)(el * P + ij);
  const bool act = el < EPB && eb + el < e1;
  const rsrc_t rmap = mkres(map + eb * P);
  const uint32_t map_ps = (uint32_t)(plane * 4);
  const uint32_t g_ps = (uint32_t)(plane * sizeof(TG));
  int32_t gid[N];
  A g6[HOIST ? N : 1][6];  // geometry of the column, kept across shots
  TriCol<A> tc;
  if constexpr (TRI) {
    for (int q = t; q < EPB * kCoef; q += blockDim.x)
      lc[q] = eb + q / kCoef < e1 ? coef[eb * kCoef + q] : TG(0);
    __syncthreads();
    if (act) {
      tri_setup<A, TG>(lc + el * kCoef, (A)glxw[i], (A)glxw[j],
                       (A)glxw[N + i] * (A)glxw[N + j], true, tc);
    }
  }
  if (act) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      gid[k] = bld<int32_t>(rmap, lo * 4u, k * map_ps);
      if constexpr (HOIST) {
        if constexpr (TRI) {
          tri_tensor<A>(tc, (A)glxw[k], (A)glxw[N + k], g6[k]);
        } else {
          const rsrc_t rg = mkres(G + (k * 6) * plane + eb * P);
#pragma unroll
          for (int c = 0; c < 6; ++c)
            g6[k][c] = (A)bld<TG>(rg, lo * (uint32_t)sizeof(TG), c * g_ps);
        }
      }
    }
  }
  const rsrc_t rs = mkres(sdiag);
  T *su = lu + el * NN, *sv = lv + el * NN;
  // shots share the nodes: the next shot's values are gathered while the
  // current one is computed
  T un[N];
  if (act) {
    const rsrc_t ru = mkres(U);
#pragma unroll
    for (int k = 0; k < N; ++k)
      un[k] = bld<T>(ru, (uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)sizeof(T), 0u);
  }
  for (int s = 0; s < nshots; ++s) {
    const rsrc_t ry = mkres(Y + (int64_t)s * N_nodes);
    T u[N];
    if (act) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        u[k] = un[k];
        su[k * P + ij] = u[k];
      }
      if (s + 1 < nshots) {
        const rsrc_t ru = mkres(U + (int64_t)(s + 1) * N_nodes);
#pragma unroll
        for (int k = 0; k < N; ++k)
          un[k] = bld<T>(ru, (uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)sizeof(T), 0u);
      }
    }
    __syncthreads();
    T v1[N], v2[N], v3[N];
    if (act) {
      T Di[N], Dj[N];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        Di[l] = Dm[i * N + l];
        Dj[l] = Dm[j * N + l];
      }
#pragma unroll
      for (int k = 0; k < N; ++k) {
        T a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          a1 += Di[l] * su[k * P + j * N + l];
          a2 += Dj[l] * su[k * P + l * N + i];
          a3 += Dm[k * N + l] * u[l];
        }
        const A b1 = a1, b2 = a2, b3 = a3;
        if constexpr (TRI && !HOIST) {
          A w1, w2, w3;
          tri_apply<A>(tc, (A)glxw[k], (A)glxw[N + k], b1, b2, b3, w1, w2, w3);
          v1[k] = (T)w1;
          v2[k] = (T)w2;
          v3[k] = (T)w3;
        } else {
          A gk[6];
          if constexpr (HOIST) {
#pragma unroll
            for (int c = 0; c < 6; ++c) gk[c] = g6[HOIST ? k : 0][c];
          } else {
            const rsrc_t rg = mkres(G + (k * 6) * plane + eb * P);
#pragma unroll
            for (int c = 0; c < 6; ++c)
              gk[c] = (A)bld<TG>(rg, lo * (uint32_t)sizeof(TG), c * g_ps);
          }
          v1[k] = (T)(gk[0] * b1 + gk[1] * b2 + gk[2] * b3);
          v2[k] = (T)(gk[1] * b1 + gk[3] * b2 + gk[4] * b3);
          v3[k] = (T)(gk[2] * b1 + gk[4] * b2 + gk[5] * b3);
        }
      }
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        su[k * P + ij] = v1[k];
        sv[k * P + ij] = v2[k];
      }
    }
    // FP32 state: the column's previous values and dt^2/M are loaded before
    // the scatter loop, whose stores to Y would otherwise order every later
    // load of Y behind them (no thread of a colour pass touches another's
    // nodes, so reading early is exact).  FP64 keeps the loads in the loop:
    // the extra registers would cost more occupancy than they save.
    constexpr bool PRE = sizeof(T) == 4;
    T yk[PRE ? N : 1], sk[PRE && MODE == kLeap ? N : 1];
    if constexpr (PRE) {
      if (act) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const uint32_t off = (uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)sizeof(T);
          yk[k] = bld<T>(ry, off, 0u);
          if constexpr (MODE == kLeap) sk[k] = bld<T>(rs, off, 0u);
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
        T o = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          o += Dti[l] * su[k * P + j * N + l];
          o += Dtj[l] * sv[k * P + l * N + i];
          o += Dm[l * N + k] * v3[l];
        }
        const int32_t m = gid[k];
        const uint32_t off = (uint32_t)(m & 0x7fffffff) * (uint32_t)sizeof(T);
        const bool first = m < 0;
        const T y = PRE ? yk[PRE ? k : 0] : bld<T>(ry, off, 0u);
        T r;
        if constexpr (MODE == kLeap) {
          const T sg = PRE ? sk[PRE ? k : 0] : bld<T>(rs, off, 0u);
          r = first ? (T(2) * u[k] - y - sg * o) : (y - sg * o);
        } else {
          r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
        }
        bst<T>(r, ry, off, 0u);
        if constexpr (GRAD) {
          if (s == 0) ufw[k] = u[k];
          else gacc += (double)o * (double)ufw[k];
        }
      }
    }
    __syncthreads();
  }
  if constexpr (GRAD) {
    // per-element sum in a fixed order (deterministic), FP64
    if (t < EPB * P) gpart[t] = act ? gacc : 0.0;
    __syncthreads();
    if (act && ij == 0) {
      double sum = 0.0;
      for (int q = 0; q < P; ++q) sum += gpart[el * P + q];
      const int64_t e = eb + el;
      grad[e] -= sum / mu_e[e];
    }
  }
}

template <int N, int EPB, typename T, typename TG, int MODE, bool HOIST, bool GRAD, bool TRI>
__global__ __launch_bounds__(EPB *N *N) void k_scalar3d(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    int64_t E, int64_t e0, int64_t e1, const T *__restrict__ U, T *__restrict__ Y,
    const T *__restrict__ sdiag, T alpha, T beta, int64_t N_nodes, int nshots, int nxcd,
    double *__restrict__ grad, const double *__restrict__ mu_e, const TG *__restrict__ coef,
    const TG *__restrict__ glxw) {
  __shared__ T lu[EPB * N * N * N];
  __shared__ T lv[EPB * N * N * N];
  __shared__ double gpart[GRAD ? EPB * N * N : 1];
  __shared__ TG lc[TRI ? EPB * kCoef : 1];
  const int64_t eb = e0 + (int64_t)xcd_block(blockIdx.x, gridDim.x, nxcd) * EPB;
  scalar3d_body<N, EPB, T, TG, MODE, HOIST, GRAD, TRI>(lu, lv, gpart, lc, map, G, Dm, E, eb, e1,
                                                       U, Y, sdiag, alpha, beta, N_nodes, nshots,
                                                       grad, mu_e, coef, glxw);
}

template <int N, int EPB, typename T, typename TG, int MODE, bool GRAD = false>
__global__ __launch_bounds__(EPB *N *N) void k_scalar2d(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    int64_t E, int64_t e0, int64_t e1, const T *__restrict__ U, T *__restrict__ Y,
    const T *__restrict__ sdiag, T alpha, T beta, int64_t N_nodes, int nshots, int nxcd,
    double *__restrict__ grad = nullptr, const double *__restrict__ mu_e = nullptr) {
  constexpr int P = N * N;
  using A = typename Mix<T, TG>::A;
  __shared__ T lu[EPB * P];
  __shared__ T lv[EPB * P];
  __shared__ double gpart[GRAD ? EPB * P : 1];
  T ufw = 0;
  double gacc = 0.0;
  const int t = threadIdx.x;
  const int el = t / P, q = t - el * P;
  const int i = q % N, j = q / N;
  const int64_t eb = e0 + (int64_t)xcd_block(blockIdx.x, gridDim.x, nxcd) * EPB;
  const int64_t plane = E * P;
  const uint32_t lo = (uint32_t)(el * P + q);
  const bool act = eb + el < e1;
  T Di[N], Dj[N], Dti[N], Dtj[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    Di[l] = Dm[i * N + l];
    Dj[l] = Dm[j * N + l];
    Dti[l] = Dm[l * N + i];
    Dtj[l] = Dm[l * N + j];
  }
  int32_t gid = 0;
  A g3[3] = {0, 0, 0};
  if (act) {
    gid = bld<int32_t>(mkres(map + eb * P), lo * 4u, 0u);
    const rsrc_t rg = mkres(G + eb * P);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      g3[c] = (A)bld<TG>(rg, lo * (uint32_t)sizeof(TG), c * (uint32_t)(plane * sizeof(TG)));
  }
  T *su = lu + el * P, *sv = lv + el * P;
  const uint32_t off = (uint32_t)(gid & 0x7fffffff) * (uint32_t)sizeof(T);
  const bool first = gid < 0;
  const rsrc_t rs = mkres(sdiag);
  for (int s = 0; s < nshots; ++s) {
    const rsrc_t ru = mkres(U + (int64_t)s * N_nodes);
    const rsrc_t ry = mkres(Y + (int64_t)s * N_nodes);
    // this node's previous value and dt^2/M with U: in flight during the pass
    T u = 0, y = 0, sg = 0;
    if (act) {
      u = bld<T>(ru, off, 0u);
      y = bld<T>(ry, off, 0u);
      if (MODE == kLeap) sg = bld<T>(rs, off, 0u);
      su[q] = u;
    }
    __syncthreads();
    T v1 = 0, v2 = 0;
    if (act) {
      T a1 = 0, a2 = 0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        a1 += Di[l] * su[j * N + l];
        a2 += Dj[l] * su[l * N + i];
      }
      const A b1 = a1, b2 = a2;
      v1 = (T)(g3[0] * b1 + g3[1] * b2);
      v2 = (T)(g3[1] * b1 + g3[2] * b2);
    }
    __syncthreads();
    if (act) {
      su[q] = v1;
      sv[q] = v2;
    }
    __syncthreads();
    if (act) {
      T o = 0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        o += Dti[l] * su[j * N + l];
        o += Dtj[l] * sv[l * N + i];
      }
      T r;
      if (MODE == kLeap) {
        r = first ? (T(2) * u - y - sg * o) : (y - sg * o);
      } else {
        r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
      }
      bst<T>(r, ry, off, 0u);
      if constexpr (GRAD) {
        if (s == 0) ufw = u;
        else gacc += (double)o * (double)ufw;
      }
    }
    __syncthreads();
  }
  if constexpr (GRAD) {
    gpart[t] = act ? gacc : 0.0;
    __syncthreads();
    if (act && q == 0) {
      double sum = 0.0;
      for (int qq = 0; qq < P; ++qq) sum += gpart[el * P + qq];
      const int64_t e = eb + el;
      grad[e] -= sum / mu_e[e];
    }
  }
}

// -------------------------------------------------------- persistent column
// One shot, n >= 5.  The grid holds only as many blocks as fit on the chip;
// each block walks a contiguous run of groups of EPB elements and issues the
// element-map and state gathers of the next group while it computes the
// current one, so the two dependent memory round trips of a group (map, then
// U) overlap the arithmetic instead of being paid once per resident block.
// Same operand order as the column kernel (bitwise equal results).  LDS:
// lu = staged U, l1 / l2 = the xi_1 / xi_2 fluxes (separate arrays so a
// group needs two barriers), lc = trilinear coefficients, sD = D.
template <int N, int EPB, typename T, typename TG, int MODE, bool TRI>
__global__ __launch_bounds__(EPB *N *N) void k_pipe3d(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    int64_t E, int64_t e0, int64_t e1, const T *__restrict__ U, T *__restrict__ Y,
    const T *__restrict__ sdiag, T alpha, T beta, int nxcd, const TG *__restrict__ coef,
    const TG *__restrict__ glxw) {
  constexpr int P = N * N, NN = P * N, NT = EPB * P;
  static_assert(EPB * kCoef <= NT, "one coefficient per thread");
  using A = TG;
  __shared__ T lu[EPB * NN], l1[EPB * NN], l2[EPB * NN];
  __shared__ T sD[N * N];
  __shared__ TG lc[TRI ? EPB * kCoef : 1];
  const int t = threadIdx.x;
  const int el = t / P, ij = t - el * P;
  const int i = ij % N, j = ij / N;
  // contiguous, balanced run of groups per block; XCD-aware block order
  const int64_t ngrp = (e1 - e0 + EPB - 1) / EPB;
  const int64_t nb = gridDim.x, lb = xcd_block(blockIdx.x, gridDim.x, nxcd);
  const int64_t q = ngrp / nb, r = ngrp % nb;
  const int64_t g0 = lb * q + (lb < r ? lb : r), g1 = g0 + q + (lb < r ? 1 : 0);
  const int64_t plane = E * P;
  const uint32_t lo = (uint32_t)(el * P + ij);
  const uint32_t map_ps = (uint32_t)(plane * 4);
  const uint32_t g_ps = (uint32_t)(plane * sizeof(TG));
  const rsrc_t ru = mkres(U), ry = mkres(Y), rs = mkres(sdiag);
  for (int qq = t; qq < N * N; qq += NT) sD[qq] = Dm[qq];
  int32_t gid[N], gidn[N];
  T u[N], un[N];
  TG cn = 0;
  if (g0 < g1) {
    const int64_t eb = e0 + g0 * EPB;
    if (eb + el < e1) {
      const rsrc_t rm = mkres(map + eb * P);
#pragma unroll
      for (int k = 0; k < N; ++k) gidn[k] = bld<int32_t>(rm, lo * 4u, k * map_ps);
#pragma unroll
      for (int k = 0; k < N; ++k)
        un[k] = bld<T>(ru, (uint32_t)(gidn[k] & 0x7fffffff) * (uint32_t)sizeof(T), 0u);
    }
    if constexpr (TRI)
      if (t < EPB * kCoef) cn = eb + t / kCoef < e1 ? coef[eb * kCoef + t] : TG(0);
  }
  for (int64_t g = g0; g < g1; ++g) {
    const int64_t eb = e0 + g * EPB;
    const bool act = eb + el < e1;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      gid[k] = gidn[k];
      u[k] = un[k];
    }
    T *su = lu + el * NN;
    if (act) {
#pragma unroll
      for (int k = 0; k < N; ++k) su[k * P + ij] = u[k];
    }
    if constexpr (TRI)
      if (t < EPB * kCoef) lc[t] = cn;
    // this group's previous values and dt^2/M (read before any store to Y)
    T yk[N], sk[MODE == kLeap ? N : 1];
    if (act) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const uint32_t off = (uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)sizeof(T);
        yk[k] = bld<T>(ry, off, 0u);
        if constexpr (MODE == kLeap) sk[k] = bld<T>(rs, off, 0u);
      }
    }
    // next group's element map (and coefficients)
    const bool nxt = g + 1 < g1;
    const int64_t ebn = eb + EPB;
    const bool actn = nxt && ebn + el < e1;
    if (actn) {
      const rsrc_t rm = mkres(map + ebn * P);
#pragma unroll
      for (int k = 0; k < N; ++k) gidn[k] = bld<int32_t>(rm, lo * 4u, k * map_ps);
    }
    if constexpr (TRI)
      if (nxt && t < EPB * kCoef) cn = ebn + t / kCoef < e1 ? coef[ebn * kCoef + t] : TG(0);
    __syncthreads();
    T v3[N];
    if (act) {
      TriCol<A> tc;
      if constexpr (TRI) {
        tri_setup<A, TG>(lc + el * kCoef, (A)glxw[i], (A)glxw[j],
                         (A)glxw[N + i] * (A)glxw[N + j], true, tc);
      }
      T *s1 = l1 + el * NN, *s2 = l2 + el * NN;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        T a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          a1 += sD[i * N + l] * su[k * P + j * N + l];
          a2 += sD[j * N + l] * su[k * P + l * N + i];
          a3 += sD[k * N + l] * u[l];
        }
        const A b1 = a1, b2 = a2, b3 = a3;
        if constexpr (TRI) {
          A w1, w2, w3;
          tri_apply<A>(tc, (A)glxw[k], (A)glxw[N + k], b1, b2, b3, w1, w2, w3);
          s1[k * P + ij] = (T)w1;
          s2[k * P + ij] = (T)w2;
          v3[k] = (T)w3;
        } else {
          A gk[6];
          const rsrc_t rg = mkres(G + (k * 6) * plane + eb * P);
#pragma unroll
          for (int c = 0; c < 6; ++c) gk[c] = (A)bld<TG>(rg, lo * (uint32_t)sizeof(TG), c * g_ps);
          s1[k * P + ij] = (T)(gk[0] * b1 + gk[1] * b2 + gk[2] * b3);
          s2[k * P + ij] = (T)(gk[1] * b1 + gk[3] * b2 + gk[4] * b3);
          v3[k] = (T)(gk[2] * b1 + gk[4] * b2 + gk[5] * b3);
        }
      }
    }
    // next group's state gathers
    if (actn) {
#pragma unroll
      for (int k = 0; k < N; ++k)
        un[k] = bld<T>(ru, (uint32_t)(gidn[k] & 0x7fffffff) * (uint32_t)sizeof(T), 0u);
    }
    __syncthreads();
    if (act) {
      const T *s1 = l1 + el * NN, *s2 = l2 + el * NN;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        T o = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          o += sD[l * N + i] * s1[k * P + j * N + l];
          o += sD[l * N + j] * s2[k * P + l * N + i];
          o += sD[l * N + k] * v3[l];
        }
        const int32_t m = gid[k];
        const uint32_t off = (uint32_t)(m & 0x7fffffff) * (uint32_t)sizeof(T);
        const bool first = m < 0;
        const T y = yk[k];
        T rr;
        if constexpr (MODE == kLeap)
          rr = first ? (T(2) * u[k] - y - sk[k] * o) : (y - sk[k] * o);
        else
          rr = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
        bst<T>(rr, ry, off, 0u);
      }
    }
  }
}

// ------------------------------------------------------- thread per node
// 3-D scalar pass with one thread per element node (EPB elements, EPB*n^3
// threads per block): every node's loads are independent, so a colour pass
// has n times more waves in flight than the column kernel to cover the
// latency of the gathers.  Threads are ordered (k, element, i + n j) so the
// map and geometry planes are read as EPB*n^2 contiguous values per k.
// The contractions use the same operand order as the column kernel.
template <int N, int EPB, typename T, typename TG, int MODE, bool GRAD, bool TRI>
__global__ __launch_bounds__(EPB *N *N *N) void k_node3d(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    int64_t E, int64_t e0, int64_t e1, const T *__restrict__ U, T *__restrict__ Y,
    const T *__restrict__ sdiag, T alpha, T beta, int64_t N_nodes, int nshots, int nxcd,
    double *__restrict__ grad, const double *__restrict__ mu_e, const TG *__restrict__ coef,
    const TG *__restrict__ glxw) {
  constexpr int P = N * N, NN = P * N, NT = EPB * NN;
  using A = TG;
  __shared__ T lu[EPB * NN], l1[EPB * NN], l2[EPB * NN], l3[EPB * NN];
  __shared__ T sD[N * N];
  __shared__ TG lc[TRI ? EPB * kCoef : 1];
  __shared__ double gpart[GRAD ? NT : 1];
  const int t = threadIdx.x;
  const int k = t / (EPB * P), rem = t - k * (EPB * P), el = rem / P, ij = rem - el * P;
  const int i = ij % N, j = ij / N;
  const int64_t eb = e0 + (int64_t)xcd_block(blockIdx.x, gridDim.x, nxcd) * EPB;
  const bool act = eb + el < e1;
  const int64_t plane = E * P;
  const uint32_t lo = (uint32_t)(el * P + ij);
  for (int q = t; q < N * N; q += NT) sD[q] = Dm[q];
  if constexpr (TRI) {
    for (int q = t; q < EPB * kCoef; q += NT)
      lc[q] = eb + q / kCoef < e1 ? coef[eb * kCoef + q] : TG(0);
  }
  int32_t gid = 0;
  A g[6] = {0, 0, 0, 0, 0, 0};
  TriCol<A> tc;
  if (act) {
    gid = bld<int32_t>(mkres(map + eb * P), lo * 4u, (uint32_t)(k * plane * 4));
    if constexpr (!TRI) {
      const rsrc_t rg = mkres(G + (k * 6) * plane + eb * P);
      const uint32_t g_ps = (uint32_t)(plane * sizeof(TG));
#pragma unroll
      for (int c = 0; c < 6; ++c) g[c] = (A)bld<TG>(rg, lo * (uint32_t)sizeof(TG), c * g_ps);
    }
  }
  __syncthreads();
  if constexpr (TRI) {
    if (act)
      tri_setup<A, TG>(lc + el * kCoef, (A)glxw[i], (A)glxw[j],
                       (A)glxw[N + i] * (A)glxw[N + j], true, tc);
  }
  const uint32_t off = (uint32_t)(gid & 0x7fffffff) * (uint32_t)sizeof(T);
  const bool first = gid < 0;
  const rsrc_t rs = mkres(sdiag);
  const int lq = el * NN + k * P + ij;
  const T *su = lu + el * NN;
  const T *s1 = l1 + el * NN, *s2 = l2 + el * NN, *s3 = l3 + el * NN;
  T ufw = 0;
  double gacc = 0.0;
  for (int s = 0; s < nshots; ++s) {
    const rsrc_t ru = mkres(U + (int64_t)s * N_nodes);
    const rsrc_t ry = mkres(Y + (int64_t)s * N_nodes);
    T u = 0, y = 0, sg = 0;
    if (act) {
      u = bld<T>(ru, off, 0u);
      y = bld<T>(ry, off, 0u);
      if constexpr (MODE == kLeap) sg = bld<T>(rs, off, 0u);
      lu[lq] = u;
    }
    __syncthreads();
    if (act) {
      T a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        a1 += sD[i * N + l] * su[k * P + j * N + l];
        a2 += sD[j * N + l] * su[k * P + l * N + i];
        a3 += sD[k * N + l] * su[l * P + ij];
      }
      const A b1 = a1, b2 = a2, b3 = a3;
      if constexpr (TRI) {
        A w1, w2, w3;
        tri_apply<A>(tc, (A)glxw[k], (A)glxw[N + k], b1, b2, b3, w1, w2, w3);
        l1[lq] = (T)w1;
        l2[lq] = (T)w2;
        l3[lq] = (T)w3;
      } else {
        l1[lq] = (T)(g[0] * b1 + g[1] * b2 + g[2] * b3);
        l2[lq] = (T)(g[1] * b1 + g[3] * b2 + g[4] * b3);
        l3[lq] = (T)(g[2] * b1 + g[4] * b2 + g[5] * b3);
      }
    }
    __syncthreads();
    if (act) {
      T o = 0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        o += sD[l * N + i] * s1[k * P + j * N + l];
        o += sD[l * N + j] * s2[k * P + l * N + i];
        o += sD[l * N + k] * s3[l * P + ij];
      }
      T r;
      if constexpr (MODE == kLeap)
        r = first ? (T(2) * u - y - sg * o) : (y - sg * o);
      else
        r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
      bst<T>(r, ry, off, 0u);
      if constexpr (GRAD) {
        if (s == 0) ufw = u;
        else gacc += (double)o * (double)ufw;
      }
    }
  }
  if constexpr (GRAD) {
    // per-element sum in a fixed order (deterministic), FP64
    gpart[t] = act ? gacc : 0.0;
    __syncthreads();
    if (act && k == 0 && ij == 0) {
      double sum = 0.0;
      for (int kk = 0; kk < N; ++kk)
        for (int q = 0; q < P; ++q) sum += gpart[kk * EPB * P + el * P + q];
      const int64_t e = eb + el;
      grad[e] -= sum / mu_e[e];
    }
  }
}

// ------------------------------------------------------------------ batch
// Many independent 3-D acoustic problems (any mesh size, orders 1..7) in one
// launch per colour: block b takes work item b = (problem, first element of
// its block); the element order selects the compiled body (blocks of 256
// threads, floor(256 / n^2) elements per block).  parity selects which of
// the problem's two state buffers holds U^{n+1} (read) and U^n (overwritten
// by U^{n+2}), as in the single-problem step.
constexpr int kBatchThreads = 256;
template <int N>
struct BatchEpb {
  static constexpr int value = kBatchThreads / (N * N);
};

template <typename T, typename TG>
__global__ __launch_bounds__(kBatchThreads) void k_batch3d(const BatchProb *__restrict__ probs,
                                                           const BatchItem *__restrict__ items,
                                                           int parity) {
  constexpr int kLds = 2 * 4 * 512;  // max over orders of 2*EPB*n^3
  constexpr int kLdsC = (kBatchThreads / 4) * kCoef;  // max EPB (order 1) * kCoef
  __shared__ T lds[kLds];
  __shared__ TG ldc[kLdsC];
  const BatchItem it = items[blockIdx.x];
  const BatchProb &p = probs[it.prob];
  const T *U = static_cast<const T *>(p.buf[parity ^ 1]);
  T *Y = static_cast<T *>(p.buf[parity]);
  const TG *G = static_cast<const TG *>(p.G);
  const T *D = static_cast<const T *>(p.Dm);
  const T *s = static_cast<const T *>(p.sdiag);
  const TG *cf = static_cast<const TG *>(p.coef);
  const TG *gw = static_cast<const TG *>(p.glxw);
#define MJW_BATCH_CASE(ORD)                                                                    \
  case ORD: {                                                                                 \
    constexpr int N = ORD + 1, EPB = BatchEpb<N>::value;                                      \
    if (p.geo == kGeoTrilinear)                                                               \
      scalar3d_body<N, EPB, T, TG, kLeap, false, false, true>(                                \
          lds, lds + EPB * N * N * N, nullptr, ldc, p.map, G, D, p.E, it.eb, it.e1, U, Y, s,  \
          T(0), T(0), p.N, 1, nullptr, nullptr, cf, gw);                                      \
    else                                                                                      \
      scalar3d_body<N, EPB, T, TG, kLeap, false, false, false>(                               \
          lds, lds + EPB * N * N * N, nullptr, ldc, p.map, G, D, p.E, it.eb, it.e1, U, Y, s,  \
          T(0), T(0), p.N, 1, nullptr, nullptr, cf, gw);                                      \
  } break;
  switch (p.order) {
    MJW_BATCH_CASE(1)
    MJW_BATCH_CASE(2)
    MJW_BATCH_CASE(3)
    MJW_BATCH_CASE(4)
    MJW_BATCH_CASE(5)
    MJW_BATCH_CASE(6)
    MJW_BATCH_CASE(7)
    default: break;
  }
#undef MJW_BATCH_CASE
}

int batch_epb(int order) {
  const int n = order + 1;
  return kBatchThreads / (n * n);
}

int launch_batch3d(int state, int geom, const BatchProb *probs, const BatchItem *items,
                   int nitems, int parity, void *stream) {
  if (nitems <= 0) return 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (state == kF64)
    hipLaunchKernelGGL((k_batch3d<double, double>), dim3(nitems), dim3(kBatchThreads), 0, st,
                       probs, items, parity);
  else if (geom == kG64)
    hipLaunchKernelGGL((k_batch3d<float, double>), dim3(nitems), dim3(kBatchThreads), 0, st,
                       probs, items, parity);
  else
    hipLaunchKernelGGL((k_batch3d<float, float>), dim3(nitems), dim3(kBatchThreads), 0, st,
                       probs, items, parity);
  return (int)hipGetLastError();
}

// Grouped batch launches: one launch per (colour, element order) over that
// order's work items, with the block size of the tile (about 128, 256 or 512
// threads) and LDS sized for that order alone, so small orders are not held to
// the largest order's LDS and registers.
template <int N, int EPB, typename T, typename TG>
__global__ __launch_bounds__(EPB *N *N) void k_batch3d_n(const BatchProb *__restrict__ probs,
                                                          const BatchItem *__restrict__ items,
                                                          int parity) {
  __shared__ T lu[EPB * N * N * N];
  __shared__ T lv[EPB * N * N * N];
  __shared__ TG lc[EPB * kCoef];
  const BatchItem it = items[blockIdx.x];
  const BatchProb &p = probs[it.prob];
  const T *U = static_cast<const T *>(p.buf[parity ^ 1]);
  T *Y = static_cast<T *>(p.buf[parity]);
  const TG *G = static_cast<const TG *>(p.G);
  const T *D = static_cast<const T *>(p.Dm);
  const T *s = static_cast<const T *>(p.sdiag);
  const TG *cf = static_cast<const TG *>(p.coef);
  const TG *gw = static_cast<const TG *>(p.glxw);
  if (p.geo == kGeoTrilinear)
    scalar3d_body<N, EPB, T, TG, kLeap, false, false, true>(lu, lv, nullptr, lc, p.map, G, D, p.E,
                                                            it.eb, it.e1, U, Y, s, T(0), T(0),
                                                            p.N, 1, nullptr, nullptr, cf, gw);
  else
    scalar3d_body<N, EPB, T, TG, kLeap, false, false, false>(lu, lv, nullptr, lc, p.map, G, D,
                                                             p.E, it.eb, it.e1, U, Y, s, T(0),
                                                             T(0), p.N, 1, nullptr, nullptr, cf,
                                                             gw);
}

// elements per block of a grouped batch tile: 0 -> 256 threads, 1 -> 512, 2 -> 128
constexpr int batch_tile_epb(int n, int tile) {
  return (tile == 1 ? 512 : tile == 2 ? 128 : 256) / (n * n) > 0
             ? (tile == 1 ? 512 : tile == 2 ? 128 : 256) / (n * n)
             : 1;
}

int batch_group_epb(int order, int tile) { return batch_tile_epb(order + 1, tile); }

template <int N, int TILE, typename T, typename TG>
static void launch_batch_n(const BatchProb *probs, const BatchItem *items, int nitems, int parity,
                           hipStream_t st) {
  constexpr int EPB = batch_tile_epb(N, TILE);
  hipLaunchKernelGGL((k_batch3d_n<N, EPB, T, TG>), dim3(nitems), dim3(EPB * N * N), 0, st, probs,
                     items, parity);
}

template <int N, typename T, typename TG>
static void launch_batch_t(int tile, const BatchProb *probs, const BatchItem *items, int nitems,
                           int parity, hipStream_t st) {
  if (tile == 1) launch_batch_n<N, 1, T, TG>(probs, items, nitems, parity, st);
  else if (tile == 2) launch_batch_n<N, 2, T, TG>(probs, items, nitems, parity, st);
  else launch_batch_n<N, 0, T, TG>(probs, items, nitems, parity, st);
}

template <int N>
static void launch_batch_g(int state, int geom, int tile, const BatchProb *probs,
                           const BatchItem *items, int nitems, int parity, hipStream_t st) {
  if (state == kF64) launch_batch_t<N, double, double>(tile, probs, items, nitems, parity, st);
  else if (geom == kG64) launch_batch_t<N, float, double>(tile, probs, items, nitems, parity, st);
  else launch_batch_t<N, float, float>(tile, probs, items, nitems, parity, st);
}

int launch_batch3d_order(int state, int geom, int order, int tile, const BatchProb *probs,
                         const BatchItem *items, int nitems, int parity, void *stream) {
  if (nitems <= 0) return 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (order + 1) {
    case 2: launch_batch_g<2>(state, geom, tile, probs, items, nitems, parity, st); break;
    case 3: launch_batch_g<3>(state, geom, tile, probs, items, nitems, parity, st); break;
    case 4: launch_batch_g<4>(state, geom, tile, probs, items, nitems, parity, st); break;
    case 5: launch_batch_g<5>(state, geom, tile, probs, items, nitems, parity, st); break;
    case 6: launch_batch_g<6>(state, geom, tile, probs, items, nitems, parity, st); break;
    case 7: launch_batch_g<7>(state, geom, tile, probs, items, nitems, parity, st); break;
    case 8: launch_batch_g<8>(state, geom, tile, probs, items, nitems, parity, st); break;
    default: return -1000;
  }
  return (int)hipGetLastError();
}

// --------------------------------------------------------------- dispatch
template <int N, int EPB, typename T, typename TG, int MODE, bool HOIST, bool GRAD, bool TRI>
static void launch3(const StiffArgs &a, hipStream_t st, unsigned blocks) {
  hipLaunchKernelGGL((k_scalar3d<N, EPB, T, TG, MODE, HOIST, GRAD, TRI>), dim3(blocks),
                     dim3(EPB * N * N), 0, st, a.map, static_cast<const TG *>(a.G),
                     static_cast<const T *>(a.Dm), a.E, a.e0, a.e1, static_cast<const T *>(a.U),
                     static_cast<T *>(a.Y), static_cast<const T *>(a.sdiag), (T)a.alpha,
                     (T)a.beta, a.N, a.nshots, a.nxcd, a.grad, a.mu,
                     static_cast<const TG *>(a.coef), static_cast<const TG *>(a.glxw));
}

// several shots in the consecutive layout (the adjoint's pair of states)
// keep the column's stored tensor in registers (HOIST); 3-D problems with
// several shots take the interleaved kernel (mjw_kernels_shots.hip)
template <int N, int EPB, typename T, typename TG, bool TRI>
static void run3(const StiffArgs &a, hipStream_t st, unsigned blocks) {
  if (a.grad)
    launch3<N, EPB, T, TG, kLeap, !TRI, true, TRI>(a, st, blocks);
  else if (a.mode == kLeap && a.nshots > 1 && !TRI)
    launch3<N, EPB, T, TG, kLeap, true, false, false>(a, st, blocks);
  else if (a.mode == kLeap)
    launch3<N, EPB, T, TG, kLeap, false, false, TRI>(a, st, blocks);
  else
    launch3<N, EPB, T, TG, kApply, false, false, TRI>(a, st, blocks);
}

template <int N, int EPB, typename T, typename TG>
static int run_scalar(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  const unsigned blocks = (unsigned)((ne + EPB - 1) / EPB);
  const unsigned threads = EPB * N * N;
  const T *D = static_cast<const T *>(a.Dm);
  const T *U = static_cast<const T *>(a.U);
  T *Y = static_cast<T *>(a.Y);
  const T *s = static_cast<const T *>(a.sdiag);
  const TG *G = static_cast<const TG *>(a.G);
  const T al = (T)a.alpha, be = (T)a.beta;
  if (a.grad && (a.mode != kLeap || a.nshots != 2)) return -1001;
  if (a.dim == 3) {
    if (a.geo == kGeoTrilinear) {
      if (!a.coef || !a.glxw) return -1002;
      run3<N, EPB, T, TG, true>(a, st, blocks);
    } else {
      run3<N, EPB, T, TG, false>(a, st, blocks);
    }
  } else if (a.grad) {
    hipLaunchKernelGGL((k_scalar2d<N, EPB, T, TG, kLeap, true>), dim3(blocks), dim3(threads), 0,
                       st, a.map, G, D, a.E, a.e0, a.e1, U, Y, s, al, be, a.N, a.nshots, a.nxcd,
                       a.grad, a.mu);
  } else if (a.mode == kLeap) {
    hipLaunchKernelGGL((k_scalar2d<N, EPB, T, TG, kLeap>), dim3(blocks), dim3(threads), 0, st,
                       a.map, G, D, a.E, a.e0, a.e1, U, Y, s, al, be, a.N, a.nshots, a.nxcd);
  } else {
    hipLaunchKernelGGL((k_scalar2d<N, EPB, T, TG, kApply>), dim3(blocks), dim3(threads), 0, st,
                       a.map, G, D, a.E, a.e0, a.e1, U, Y, s, al, be, a.N, a.nshots, a.nxcd);
  }
  return (int)hipGetLastError();
}

template <int N, int EPB, typename T, typename TG, int MODE, bool GRAD, bool TRI>
static void launch_node(const StiffArgs &a, hipStream_t st, unsigned blocks) {
  hipLaunchKernelGGL((k_node3d<N, EPB, T, TG, MODE, GRAD, TRI>), dim3(blocks),
                     dim3(EPB * N * N * N), 0, st, a.map, static_cast<const TG *>(a.G),
                     static_cast<const T *>(a.Dm), a.E, a.e0, a.e1, static_cast<const T *>(a.U),
                     static_cast<T *>(a.Y), static_cast<const T *>(a.sdiag), (T)a.alpha,
                     (T)a.beta, a.N, a.nshots, a.nxcd, a.grad, a.mu,
                     static_cast<const TG *>(a.coef), static_cast<const TG *>(a.glxw));
}

template <int N, int EPB, typename T, typename TG, bool TRI>
static void run_node3(const StiffArgs &a, hipStream_t st, unsigned blocks) {
  if (a.grad)
    launch_node<N, EPB, T, TG, kLeap, true, TRI>(a, st, blocks);
  else if (a.mode == kLeap)
    launch_node<N, EPB, T, TG, kLeap, false, TRI>(a, st, blocks);
  else
    launch_node<N, EPB, T, TG, kApply, false, TRI>(a, st, blocks);
}

template <int N, int EPB, typename T, typename TG>
static int run_node(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  if (a.dim != 3 || a.ncomp != 1) return -1003;
  if (a.grad && (a.mode != kLeap || a.nshots != 2)) return -1001;
  const unsigned blocks = (unsigned)((ne + EPB - 1) / EPB);
  if (a.geo == kGeoTrilinear) {
    if (!a.coef || !a.glxw) return -1002;
    run_node3<N, EPB, T, TG, true>(a, st, blocks);
  } else {
    run_node3<N, EPB, T, TG, false>(a, st, blocks);
  }
  return (int)hipGetLastError();
}

template <int N, int EPB, typename T, typename TG, int MODE, bool TRI>
static int launch_pipe(const StiffArgs &a, hipStream_t st, int64_t ngrp) {
  // resident blocks on the device (cached per kernel and device)
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1004;
  if (!cached[dev]) {
    int per_cu = 0, ncu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(&k_pipe3d<N, EPB, T, TG, MODE, TRI>),
            EPB * N * N, 0) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return -1004;
    cached[dev] = std::max(1, per_cu) * std::max(1, ncu);
  }
  const unsigned blocks = (unsigned)std::min<int64_t>(ngrp, cached[dev]);
  hipLaunchKernelGGL((k_pipe3d<N, EPB, T, TG, MODE, TRI>), dim3(blocks), dim3(EPB * N * N), 0, st,
                     a.map, static_cast<const TG *>(a.G), static_cast<const T *>(a.Dm), a.E, a.e0,
                     a.e1, static_cast<const T *>(a.U), static_cast<T *>(a.Y),
                     static_cast<const T *>(a.sdiag), (T)a.alpha, (T)a.beta, a.nxcd,
                     static_cast<const TG *>(a.coef), static_cast<const TG *>(a.glxw));
  return (int)hipGetLastError();
}

// persistent column kernel for one shot; other cases (several shots, the
// adjoint step) take the column kernel with the same tile
template <int N, int EPB, typename T, typename TG>
static int run_pipe(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  if (a.dim != 3 || a.ncomp != 1) return -1003;
  if (a.grad || a.nshots != 1) return run_scalar<N, EPB, T, TG>(a, st);
  const int64_t ngrp = (ne + EPB - 1) / EPB;
  const bool tri = a.geo == kGeoTrilinear;
  if (tri && (!a.coef || !a.glxw)) return -1002;
  if (a.mode == kLeap)
    return tri ? launch_pipe<N, EPB, T, TG, kLeap, true>(a, st, ngrp)
               : launch_pipe<N, EPB, T, TG, kLeap, false>(a, st, ngrp);
  return tri ? launch_pipe<N, EPB, T, TG, kApply, true>(a, st, ngrp)
             : launch_pipe<N, EPB, T, TG, kApply, false>(a, st, ngrp);
}

template <int N, int EPB>
static int run_pipe_prec(const StiffArgs &a, hipStream_t st) {
  if (a.state == kF64) return run_pipe<N, EPB, double, double>(a, st);
  if (a.geom == kG64) return run_pipe<N, EPB, float, double>(a, st);
  return run_pipe<N, EPB, float, float>(a, st);
}

template <int N, int EPB>
static int run_node_prec(const StiffArgs &a, hipStream_t st) {
  if (a.state == kF64) return run_node<N, EPB, double, double>(a, st);
  if (a.geom == kG64) return run_node<N, EPB, float, double>(a, st);
  return run_node<N, EPB, float, float>(a, st);
}

template <int N, int EPB>
static int run_prec(const StiffArgs &a, hipStream_t st) {
  if (a.state == kF64) return run_scalar<N, EPB, double, double>(a, st);
  if (a.geom == kG64) return run_scalar<N, EPB, float, double>(a, st);
  return run_scalar<N, EPB, float, float>(a, st);
}

// tile (elements per block) candidates; the first is the default.
// 3-D: threads = EPB*n^2, LDS = 2*EPB*n^3*sizeof(T) <= 64 KiB.
#define MJW_TILES3(X) \
  X(2, 16) X(2, 32) X(3, 7) X(3, 14) X(4, 8) X(4, 4) X(4, 16) X(5, 10) X(5, 5) X(5, 20) \
  X(6, 7) X(6, 14) X(6, 16) X(7, 5) X(7, 10) X(8, 2) X(8, 4)
// thread-per-node tiles (elements per block; listed as -EPB):
// threads = EPB*n^3 <= 1024, LDS = 4*EPB*n^3*sizeof(T) + ...
#define MJW_NODE_TILES3(X) \
  X(2, 32) X(2, 64) X(3, 8) X(3, 16) X(4, 4) X(4, 8) X(5, 2) X(5, 4) X(6, 1) X(6, 2) X(6, 4) \
  X(7, 1) X(7, 2) X(8, 1) X(8, 2)
// persistent column tiles (listed as kPipeTile + EPB), n >= 5
#define MJW_PIPE_TILES3(X) X(5, 5) X(5, 10) X(6, 7) X(6, 14) X(7, 5) X(7, 10) X(8, 4)
#define MJW_TILES2(X) \
  X(2, 64) X(2, 16) X(3, 28) X(3, 7) X(4, 16) X(4, 4) X(5, 10) X(5, 20) X(6, 7) X(6, 14) \
  X(7, 5) X(7, 10) X(8, 4) X(8, 1)

int launch_stiffness_scalar(const StiffArgs &a, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n = a.order + 1;
#define X(NN, EE) \
  if (n == NN && a.epb == EE) return run_prec<NN, EE>(a, st);
#define XN(NN, EE) \
  if (n == NN && a.epb == -EE) return run_node_prec<NN, EE>(a, st);
#define XP(NN, EE) \
  if (n == NN && a.epb == kPipeTile + EE) return run_pipe_prec<NN, EE>(a, st);
  if (a.dim == 3) {
    MJW_TILES3(X)
    MJW_NODE_TILES3(XN)
    MJW_PIPE_TILES3(XP)
  } else {
    MJW_TILES2(X)
  }
#undef X
#undef XN
#undef XP
  return -1000;
}

std::vector<int> tile_choices_scalar(int dim, int order) {
  std::vector<int> v;
  const int n = order + 1;
#define X(NN, EE) \
  if (n == NN) v.push_back(EE);
#define XN(NN, EE) \
  if (n == NN) v.push_back(-EE);
#define XP(NN, EE) \
  if (n == NN) v.push_back(kPipeTile + EE);
  if (dim == 3) {
    MJW_TILES3(X)
    MJW_NODE_TILES3(XN)
    MJW_PIPE_TILES3(XP)
  } else {
    MJW_TILES2(X)
  }
#undef X
#undef XN
#undef XP
  return v;
}

}  // namespace mjw
