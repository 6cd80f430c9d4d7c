// Stiffness passes of 3-D problems with several shots (right-hand sides) in
// the shot-interleaved layout u[(m*N + g)*S + s], gfx950: acoustic
// (k_shots3d) and isotropic elastic (k_shots_el).
//
// Same arithmetic per node and shot as the column kernel
// (mjw_kernels_scalar.hip, restating MltAdd_ElementH1_Scalar,
// Computation/ProdMatVectScalarH1.cxx:256-668), mapped for many shots: a
// block takes a run of R consecutive elements of a colour (R = 1, 2, 4) and
// ST consecutive shots; thread (i + n j, lane) owns column (i, j) of shot
// s0 + lane and loops k.  The ST lanes of a column gather one contiguous run
// of ST values per node, so a wavefront's loads of U, Y and dt^2/M are ST
// times fewer cache lines than ST separate single-shot passes.  Each
// element's geometry is built once into LDS, one thread per node, and read
// by every shot: the adjugate rows of DF and mu w / J per node (trilinear
// map; the flux is the single-shot kernels' f R (R^T b), so results equal
// single-shot runs bit for bit) or the stored tensor.  With R > 1 the next
// element's map and U are loaded while the current one is computed.  LDS
// state is [k][i + n j][lane] so lanes hit consecutive banks.
#include <hip/hip_runtime.h>

#include "mjw_device.hpp"
#include "mjw_internal.hpp"
#include "mjw_tri.hpp"

namespace mjw {

template <int N, int ST, int R, typename T, typename TG, int MODE, bool TRI>
__global__ __launch_bounds__(N *N *ST) void k_shots3d(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    int64_t E, int64_t e0, int64_t e1, const T *__restrict__ U, T *__restrict__ Y,
    const T *__restrict__ sdiag, T alpha, T beta, int nshots, int nxcd,
    const TG *__restrict__ coef, const TG *__restrict__ glxw) {
  constexpr int P = N * N, NN = P * N;
  using A = TG;
  // su holds the staged U, then (after a barrier) the xi_1 fluxes
  __shared__ T su[NN * ST], s2[NN * ST];
  T *const s1 = su;
  constexpr int NG = TRI ? 10 : 6;  // per node: r0 r1 r2 f, or the 6 tensor entries
  __shared__ A lg[P * N * NG];      // [i + n j][k][NG]
  // D and D^T: read from LDS inside the element loop (not held in registers)
  __shared__ T sD[P], sDt[P];
  const int t = threadIdx.x;
  const int lane = t % ST, ij = t / ST;
  const int i = ij % N, j = ij / N;
  for (int q = t; q < P; q += P * ST) {
    sD[q] = Dm[q];
    sDt[q] = Dm[(q % N) * N + q / N];
  }
  // this block's run of R consecutive elements of the colour and its shot
  // tile; the tiles of one run are consecutive logical blocks (same XCD, same
  // time), so they share the L2 lines of the interleaved state
  const unsigned ntile = (unsigned)((nshots + ST - 1) / ST);
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x, nxcd);
  const int64_t eb = e0 + (int64_t)(lb / ntile) * R;
  const int64_t ee = eb + R < e1 ? eb + R : e1;
  const int s = (int)(lb % ntile) * ST + lane;
  const bool sok = s < nshots;
  const rsrc_t ru = mkres(U), ry = mkres(Y), rs = mkres(sdiag);
  int32_t gid[N];
  T u[N];
  if (eb < ee) {
#pragma unroll
    for (int k = 0; k < N; ++k) gid[k] = map[((int64_t)k * E + eb) * P + ij];
    if (sok)
#pragma unroll
      for (int k = 0; k < N; ++k)
        u[k] = bld<T>(ru, ((uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)nshots + (uint32_t)s) *
                              (uint32_t)sizeof(T), 0u);
  }
  for (int64_t e = eb; e < ee; ++e) {
    const bool nxt = R > 1 && e + 1 < ee;
    // stage U; one thread per node builds the element's geometry
    if (sok)
#pragma unroll
      for (int k = 0; k < N; ++k) su[(k * P + ij) * ST + lane] = u[k];
    for (int gq = t; gq < NN; gq += P * ST) {  // node gq = i + n j + n^2 k
      const int gij = gq % P, gk = gq / P;
      A *q = lg + (gij * N + gk) * NG;
      if constexpr (TRI) {
        const TG *c = coef + e * kCoef;
        const int gi = gij % N, gj = gij / N;
        TriCol<A> tc;
        tri_setup<A, TG>(c, (A)glxw[gi], (A)glxw[gj], (A)glxw[N + gi] * (A)glxw[N + gj], true, tc);
        tri_rows_f<A>(tc, (A)glxw[gk], (A)glxw[N + gk], q, q + 3, q + 6, q[9]);
      } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) q[c] = (A)G[((int64_t)(gk * 6 + c) * E + e) * P + gij];
      }
    }
    // next element's map; this element's previous values and dt^2/M
    int32_t gn[N];
    if (nxt)
#pragma unroll
      for (int k = 0; k < N; ++k) gn[k] = map[((int64_t)k * E + e + 1) * P + ij];
    T yk[N], sk[MODE == kLeap ? N : 1];
    auto load_y = [&]() {
      if (sok) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const uint32_t g = (uint32_t)(gid[k] & 0x7fffffff);
          yk[k] = bld<T>(ry, (g * (uint32_t)nshots + (uint32_t)s) * (uint32_t)sizeof(T), 0u);
          if constexpr (MODE == kLeap) sk[k] = bld<T>(rs, g * (uint32_t)sizeof(T), 0u);
        }
      }
    };
    // runs load Y early (hidden behind the gradient); single elements load it
    // after the gradient, which keeps fewer registers live
    if constexpr (R > 1) load_y();
    __syncthreads();
    T v1[N], v2[N], v3[N];
    if (sok) {
      T Di[N], Dj[N];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        Di[l] = sD[i * N + l];
        Dj[l] = sD[j * N + l];
      }
#pragma unroll
      for (int k = 0; k < N; ++k) {
        T a1 = 0, a2 = 0, a3 = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          a1 += Di[l] * su[(k * P + j * N + l) * ST + lane];
          a2 += Dj[l] * su[(k * P + l * N + i) * ST + lane];
          a3 += Dm[k * N + l] * u[l];
        }
        const A b1 = a1, b2 = a2, b3 = a3;
        const A *g = lg + (ij * N + k) * NG;
        if constexpr (TRI) {
          A w1, w2, w3;
          tri_flux<A>(g, g + 3, g + 6, g[9], b1, b2, b3, w1, w2, w3);
          v1[k] = (T)w1;
          v2[k] = (T)w2;
          v3[k] = (T)w3;
        } else {
          v1[k] = (T)(g[0] * b1 + g[1] * b2 + g[2] * b3);
          v2[k] = (T)(g[1] * b1 + g[3] * b2 + g[4] * b3);
          v3[k] = (T)(g[2] * b1 + g[4] * b2 + g[5] * b3);
        }
      }
    }
    if constexpr (R == 1) load_y();
    // next element's U, in flight during the divergence
    T un[N];
    if (nxt && sok)
#pragma unroll
      for (int k = 0; k < N; ++k)
        un[k] = bld<T>(ru, ((uint32_t)(gn[k] & 0x7fffffff) * (uint32_t)nshots + (uint32_t)s) *
                               (uint32_t)sizeof(T), 0u);
    __syncthreads();  // every read of the staged U and the geometry is done
    if (sok) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        s1[(k * P + ij) * ST + lane] = v1[k];
        s2[(k * P + ij) * ST + lane] = v2[k];
      }
    }
    __syncthreads();
    if (sok) {
      T Dti[N], Dtj[N];
#pragma unroll
      for (int l = 0; l < N; ++l) {
        Dti[l] = sDt[i * N + l];
        Dtj[l] = sDt[j * N + l];
      }
#pragma unroll
      for (int k = 0; k < N; ++k) {
        T o = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          o += Dti[l] * s1[(k * P + j * N + l) * ST + lane];
          o += Dtj[l] * s2[(k * P + l * N + i) * ST + lane];
          o += Dm[l * N + k] * v3[l];
        }
        const int32_t m = gid[k];
        const uint32_t off =
            ((uint32_t)(m & 0x7fffffff) * (uint32_t)nshots + (uint32_t)s) * (uint32_t)sizeof(T);
        const bool first = m < 0;
        const T y = yk[k];
        T r;
        if constexpr (MODE == kLeap)
          r = first ? (T(2) * u[k] - y - sk[k] * o) : (y - sk[k] * o);
        else
          r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
        bst<T>(r, ry, off, 0u);
      }
    }
    if (nxt) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        gid[k] = gn[k];
        u[k] = un[k];
      }
      __syncthreads();  // the fluxes are read before the next element is staged
    }
  }
}

// shots per block (lanes per column): the largest power of two <= 16 with
// n^2*ST <= 1024 threads and LDS = 2*n^3*ST*sizeof(T) + 10*n^3*sizeof(TG)
// + kCoef*sizeof(TG) <= 64 KiB; fewer when the problem has fewer shots
template <int N, typename T, typename TG>
struct ShotTile {
  static constexpr int lds(int st) {
    return 2 * N * N * N * st * (int)sizeof(T) + 10 * N * N * N * (int)sizeof(TG) +
           kCoef * (int)sizeof(TG);
  }
  static constexpr int pick() {
    int st = 16;
    while (st > 1 && (lds(st) > 65536 || N * N * st > 1024)) st /= 2;
    return st;
  }
  static constexpr int value = pick();
};

template <int N, int ST, int R, typename T, typename TG>
static int run_shots_st(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  const dim3 grid((unsigned)(((ne + R - 1) / R) * ((a.nshots + ST - 1) / ST)));
  const dim3 block(N * N * ST);
  const T *D = static_cast<const T *>(a.Dm);
  const T *U = static_cast<const T *>(a.U);
  T *Y = static_cast<T *>(a.Y);
  const T *s = static_cast<const T *>(a.sdiag);
  const TG *G = static_cast<const TG *>(a.G);
  const TG *cf = static_cast<const TG *>(a.coef);
  const TG *gw = static_cast<const TG *>(a.glxw);
  const bool tri = a.geo == kGeoTrilinear;
  if (tri && (!cf || !gw)) return -1002;
#define MJW_SHOTS_LAUNCH(RR, MODE, TRI)                                                        \
  hipLaunchKernelGGL((k_shots3d<N, ST, RR, T, TG, MODE, TRI>), grid, block, 0, st, a.map, G, D, \
                     a.E, a.e0, a.e1, U, Y, s, (T)a.alpha, (T)a.beta, a.nshots, a.nxcd, cf, gw)
  if (a.mode == kLeap) {
    if (tri) MJW_SHOTS_LAUNCH(R, kLeap, true);
    else MJW_SHOTS_LAUNCH(R, kLeap, false);
  } else {
    // operator applications (setup, CFL, energy) run one element per block
    if (R != 1) return -1005;
    if (tri) MJW_SHOTS_LAUNCH(1, kApply, true);
    else MJW_SHOTS_LAUNCH(1, kApply, false);
  }
#undef MJW_SHOTS_LAUNCH
  return (int)hipGetLastError();
}

// launch shape code: ST shot lanes per column and runs of R elements per
// block, code = ST for R = 1, 100 R + ST otherwise (R = 2, 4 with ST >= 8);
// 0 picks the smallest power of two >= nshots within the limit, R = 1
template <int N, typename T, typename TG>
static int run_shots_t(const StiffArgs &a, hipStream_t st) {
  constexpr int SM = ShotTile<N, T, TG>::value;
  static_assert(SM >= 1, "shot tile");
  if (a.epb == kTileMfma) return launch_stiffness_shots_mfma(a, st);
  int want = a.epb % 100, R = a.epb / 100;
  if (a.epb <= 0) {
    want = 1;
    R = 0;
    while (want < a.nshots && want < SM) want *= 2;
  }
  if (want > SM) return -1005;
  // operator applications take the one-element form of the same lanes
  if (a.mode != kLeap) R = 0;
  if constexpr (SM >= 16)
    if (want == 16) {
      if (R == 4) return run_shots_st<N, 16, 4, T, TG>(a, st);
      if (R == 2) return run_shots_st<N, 16, 2, T, TG>(a, st);
      if (R == 0) return run_shots_st<N, 16, 1, T, TG>(a, st);
      return -1005;
    }
  if constexpr (SM >= 8)
    if (want == 8) {
      if (R == 4) return run_shots_st<N, 8, 4, T, TG>(a, st);
      if (R == 2) return run_shots_st<N, 8, 2, T, TG>(a, st);
      if (R == 0) return run_shots_st<N, 8, 1, T, TG>(a, st);
      return -1005;
    }
  if (R != 0) return -1005;
  if constexpr (SM >= 4)
    if (want == 4) return run_shots_st<N, 4, 1, T, TG>(a, st);
  if constexpr (SM >= 2)
    if (want == 2) return run_shots_st<N, 2, 1, T, TG>(a, st);
  if (want == 1) return run_shots_st<N, 1, 1, T, TG>(a, st);
  return -1005;
}

template <int N>
static void shot_tiles_n(int state, int geom, std::vector<int> &v) {
  int sm;
  if (state == kF64) sm = ShotTile<N, double, double>::value;
  else if (geom == kG64) sm = ShotTile<N, float, double>::value;
  else sm = ShotTile<N, float, float>::value;
  for (int t = sm; t >= 1; t /= 2) {
    v.push_back(t);
    if (t >= 8) {
      v.push_back(200 + t);
      v.push_back(400 + t);
    }
  }
}

std::vector<int> tile_choices_shots(int order, int state, int geom) {
  std::vector<int> v;
  switch (order + 1) {
    case 2: shot_tiles_n<2>(state, geom, v); break;
    case 3: shot_tiles_n<3>(state, geom, v); break;
    case 4: shot_tiles_n<4>(state, geom, v); break;
    case 5: shot_tiles_n<5>(state, geom, v); break;
    case 6: shot_tiles_n<6>(state, geom, v); break;
    case 7: shot_tiles_n<7>(state, geom, v); break;
    case 8: shot_tiles_n<8>(state, geom, v); break;
    default: break;
  }
  if (shots_mfma_supported(order, state)) v.push_back(kTileMfma);
  return v;
}

template <int N>
static int run_shots_n(const StiffArgs &a, hipStream_t st) {
  if (a.state == kF64) return run_shots_t<N, double, double>(a, st);
  if (a.geom == kG64) return run_shots_t<N, float, double>(a, st);
  return run_shots_t<N, float, float>(a, st);
}

int launch_stiffness_shots(const StiffArgs &a, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (a.order + 1) {
    case 2: return run_shots_n<2>(a, st);
    case 3: return run_shots_n<3>(a, st);
    case 4: return run_shots_n<4>(a, st);
    case 5: return run_shots_n<5>(a, st);
    case 6: return run_shots_n<6>(a, st);
    case 7: return run_shots_n<7>(a, st);
    case 8: return run_shots_n<8>(a, st);
    default: return -1000;
  }
}

}  // namespace mjw

namespace mjw {

// ----------------------------------------------------------------- elastic
// Isotropic elastodynamics, several shots interleaved (u[((m*N + g)*S + s]):
// the arithmetic of k_elastic3d (mjw_kernels_elastic.hip, restating
// MltAdd_ElementH1, Computation/ProdMatVectH1.cxx:674-944) with one element
// and ST shot lanes per block.  The factor M = sqrt(J w) DF^{-1} of each node
// is built once per block into LDS (trilinear map or stored planes).  LDS
// planes [6][k][i + n j][lane]: u_m (0..2); slab k of them is overwritten by
// its xi_1 fluxes once every thread has read it (one barrier per slab); xi_2
// fluxes in planes 3..5.
template <int N, int ST, typename T, typename TG, int MODE, bool TRI>
__global__ __launch_bounds__(N *N *ST) void k_shots_el(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    const double *__restrict__ lam_e, const double *__restrict__ mu_e, int64_t E, int64_t e0,
    int64_t e1, const T *__restrict__ U, T *__restrict__ Y, const T *__restrict__ sdiag,
    T alpha, T beta, int64_t N_nodes, int nshots, int nxcd, const TG *__restrict__ coef,
    const TG *__restrict__ glxw) {
  constexpr int P = N * N, NN = P * N, PL = NN * ST;
  using A = TG;
  __shared__ T lds[6 * PL];
  __shared__ A lm[P * N * 9];  // [i + n j][k][9]
  __shared__ TG lc[TRI ? kCoef : 1];
  const int t = threadIdx.x;
  const int lane = t % ST, ij = t / ST;
  const int i = ij % N, j = ij / N;
  // element and shot tile; the tiles of one element are consecutive logical
  // blocks (same XCD, same time), so they share the L2 lines of the
  // interleaved state
  const unsigned ntile = (unsigned)((nshots + ST - 1) / ST);
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x, nxcd);
  const int64_t e = e0 + (int64_t)(lb / ntile);
  const int s = (int)(lb % ntile) * ST + lane;
  const bool act = e < e1, sact = act && s < nshots;
  if constexpr (TRI) {
    for (int q = t; q < kCoef; q += P * ST)
      if (act) lc[q] = coef[e * kCoef + q];
    __syncthreads();
    // one thread per node gq = i + n j + n^2 k (the first n^3 threads)
    if (act)
      for (int gq = t; gq < NN; gq += P * ST) {
        const int gij = gq % P, gk = gq / P, gi = gij % N, gj = gij / N;
        TriCol<A> tc;
        tri_setup<A, TG>(lc, (A)glxw[gi], (A)glxw[gj], (A)glxw[N + gi] * (A)glxw[N + gj], false,
                         tc);
        tri_m<A>(tc, (A)glxw[gk], (A)glxw[N + gk], lm + (gij * N + gk) * 9);
      }
  } else {
    if (act)
      for (int gq = t; gq < NN; gq += P * ST) {
        const int gij = gq % P, gk = gq / P;
#pragma unroll
        for (int c = 0; c < 9; ++c)
          lm[(gij * N + gk) * 9 + c] = (A)G[((int64_t)(gk * 9 + c) * E + e) * P + gij];
      }
  }
  int32_t gid[N];
  A lam = 0, mu = 0;
  if (act) {
#pragma unroll
    for (int k = 0; k < N; ++k) gid[k] = map[((int64_t)k * E + e) * P + ij];
    lam = (A)lam_e[e];
    mu = (A)mu_e[e];
  }
  const rsrc_t ru = mkres(U), ry = mkres(Y), rs = mkres(sdiag);
  // byte offset of node g's shot s in component 0; component m adds the
  // scalar m * cs (the buffer's SGPR offset)
  auto slot = [&](int32_t g) {
    return ((uint32_t)(g & 0x7fffffff) * (uint32_t)nshots + (uint32_t)s) * (uint32_t)sizeof(T);
  };
  const uint32_t cs = (uint32_t)N_nodes * (uint32_t)nshots * (uint32_t)sizeof(T);
  T u[3][N];
  if (sact) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int k = 0; k < N; ++k) {
        u[m][k] = bld<T>(ru, slot(gid[k]), (uint32_t)m * cs);
        lds[m * PL + (k * P + ij) * ST + lane] = u[m][k];
      }
  }
  __syncthreads();
  // per layer k: the gradients of slab k, a barrier (every read of slab k of
  // the state planes is done), then the slab's xi_1 fluxes overwrite it
  T v2[3][N];
  T Di[N], Dj[N];
  if (sact) {
#pragma unroll
    for (int l = 0; l < N; ++l) {
      Di[l] = Dm[i * N + l];
      Dj[l] = Dm[j * N + l];
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    A g[3][3];
    if (sact) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          a0 += Di[l] * lds[m * PL + (k * P + j * N + l) * ST + lane];
          a1 += Dj[l] * lds[m * PL + (k * P + l * N + i) * ST + lane];
          a2 += Dm[k * N + l] * u[m][l];
        }
        g[m][0] = a0;
        g[m][1] = a1;
        g[m][2] = a2;
      }
    }
    __syncthreads();
    if (sact) {
      const A *M = lm + (ij * N + k) * 9;
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
        lds[m * PL + (k * P + ij) * ST + lane] =
            (T)(M[0] * sg[m][0] + M[1] * sg[m][1] + M[2] * sg[m][2]);
        lds[(3 + m) * PL + (k * P + ij) * ST + lane] =
            (T)(M[3] * sg[m][0] + M[4] * sg[m][1] + M[5] * sg[m][2]);
        v2[m][k] = (T)(M[6] * sg[m][0] + M[7] * sg[m][1] + M[8] * sg[m][2]);
      }
    }
  }
  // previous values and dt^2/M, in flight across the barrier (all read
  // before this thread's first store to Y)
  T yk[3][N], sk[MODE == kLeap ? N : 1];
  if (sact) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
      for (int m = 0; m < 3; ++m) yk[m][k] = bld<T>(ry, slot(gid[k]), (uint32_t)m * cs);
      if constexpr (MODE == kLeap)
        sk[k] = bld<T>(rs, (uint32_t)(gid[k] & 0x7fffffff) * (uint32_t)sizeof(T), 0u);
    }
  }
  __syncthreads();
  if (sact) {
    T Dti[N], Dtj[N];
#pragma unroll
    for (int l = 0; l < N; ++l) {
      Dti[l] = Dm[l * N + i];
      Dtj[l] = Dm[l * N + j];
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const int32_t mm = gid[k];
      const bool first = mm < 0;
      T sgd = 0;
      if constexpr (MODE == kLeap) sgd = sk[k];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        T o = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
          o += Dti[l] * lds[m * PL + (k * P + j * N + l) * ST + lane];
          o += Dtj[l] * lds[(3 + m) * PL + (k * P + l * N + i) * ST + lane];
          o += Dm[l * N + k] * v2[m][l];
        }
        const uint32_t off = slot(mm);
        const T y = yk[m][k];
        T r;
        if constexpr (MODE == kLeap)
          r = first ? (T(2) * u[m][k] - y - sgd * o) : (y - sgd * o);
        else
          r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
        bst<T>(r, ry, off, (uint32_t)m * cs);
      }
    }
  }
}

// The same pass with one thread per (node, shot): thread (q, lane), q = i +
// n j + n^2 k.  A block is n^3 ST threads, so for n = 5 and ST = 8 (1000
// threads) 98 % of the lanes carry work, where the column mapping (25 ST
// threads) leaves a quarter of its last wave idle.  All three contractions
// read their neighbours from LDS; after a barrier the nine flux planes
// overwrite the three state planes.  The sums run in the column kernel's
// order (a0, a1, a2 per l; xi_1, xi_2, xi_3 terms per l in the divergence), so
// both mappings give the same bits.  The component offset of an interleaved
// value is a scalar (the buffer's SGPR offset): one VGPR offset per thread.
template <int N, int ST, typename T, typename TG, int MODE, bool TRI>
__global__ __launch_bounds__(N *N *N *ST) void k_shots_el_node(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const T *__restrict__ Dm,
    const double *__restrict__ lam_e, const double *__restrict__ mu_e, int64_t E, int64_t e0,
    int64_t e1, const T *__restrict__ U, T *__restrict__ Y, const T *__restrict__ sdiag,
    T alpha, T beta, int64_t N_nodes, int nshots, int nxcd, const TG *__restrict__ coef,
    const TG *__restrict__ glxw) {
  constexpr int P = N * N, NN = P * N, PL = NN * ST;
  using A = TG;
  __shared__ T lds[9 * PL];     // u_m (0..2); then fluxes xi_1 (0..2), xi_2 (3..5), xi_3 (6..8)
  __shared__ A lm[NN * 9];      // [q][9]
  __shared__ TG lc[TRI ? kCoef : 1];
  const int t = threadIdx.x;
  const int lane = t % ST, q = t / ST;
  const int i = q % N, j = (q / N) % N, k = q / P;
  // element and shot tile; the tiles of one element are consecutive logical
  // blocks (same XCD, same time), so they share the L2 lines of the
  // interleaved state
  const unsigned ntile = (unsigned)((nshots + ST - 1) / ST);
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x, nxcd);
  const int64_t e = e0 + (int64_t)(lb / ntile);
  const int s = (int)(lb % ntile) * ST + lane;
  const bool act = e < e1, sact = act && s < nshots;
  // the element's factor M = sqrt(J w) DF^{-1} per node, built by the first
  // n^3 threads (whole waves, the others wait at the barrier)
  if constexpr (TRI) {
    for (int c = t; c < kCoef; c += NN * ST)
      if (act) lc[c] = coef[e * kCoef + c];
    __syncthreads();
    if (act && t < NN) {
      const int gij = t % P, gk = t / P, gi = gij % N, gj = gij / N;
      TriCol<A> tc;
      tri_setup<A, TG>(lc, (A)glxw[gi], (A)glxw[gj], (A)glxw[N + gi] * (A)glxw[N + gj], false, tc);
      tri_m<A>(tc, (A)glxw[gk], (A)glxw[N + gk], lm + t * 9);
    }
  } else {
    if (act && t < NN) {
      const int gij = t % P, gk = t / P;
#pragma unroll
      for (int c = 0; c < 9; ++c) lm[t * 9 + c] = (A)G[((int64_t)(gk * 9 + c) * E + e) * P + gij];
    }
  }
  int32_t gq = 0;
  A lam = 0, mu = 0;
  if (act) {
    gq = map[((int64_t)k * E + e) * P + i + N * j];
    lam = (A)lam_e[e];
    mu = (A)mu_e[e];
  }
  const rsrc_t ru = mkres(U), ry = mkres(Y), rs = mkres(sdiag);
  const uint32_t vo = ((uint32_t)(gq & 0x7fffffff) * (uint32_t)nshots + (uint32_t)s) *
                      (uint32_t)sizeof(T);
  const uint32_t cs = (uint32_t)N_nodes * (uint32_t)nshots * (uint32_t)sizeof(T);
  T u[3], yk[3], sk = 0;
  if (sact) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      u[m] = bld<T>(ru, vo, (uint32_t)m * cs);
      lds[m * PL + t] = u[m];
    }
    // previous values and dt^2/M, in flight across the barriers
#pragma unroll
    for (int m = 0; m < 3; ++m) yk[m] = bld<T>(ry, vo, (uint32_t)m * cs);
    if constexpr (MODE == kLeap)
      sk = bld<T>(rs, (uint32_t)(gq & 0x7fffffff) * (uint32_t)sizeof(T), 0u);
  }
  __syncthreads();
  T f1[3], f2[3], f3[3];
  if (sact) {
    T Di[N], Dj[N], Dk[N];
#pragma unroll
    for (int l = 0; l < N; ++l) {
      Di[l] = Dm[i * N + l];
      Dj[l] = Dm[j * N + l];
      Dk[l] = Dm[k * N + l];
    }
    A g[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        a0 += Di[l] * lds[m * PL + (l + N * j + P * k) * ST + lane];
        a1 += Dj[l] * lds[m * PL + (i + N * l + P * k) * ST + lane];
        a2 += Dk[l] * lds[m * PL + (i + N * j + P * l) * ST + lane];
      }
      g[m][0] = a0;
      g[m][1] = a1;
      g[m][2] = a2;
    }
    const A *M = lm + q * 9;
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
      f1[m] = (T)(M[0] * sg[m][0] + M[1] * sg[m][1] + M[2] * sg[m][2]);
      f2[m] = (T)(M[3] * sg[m][0] + M[4] * sg[m][1] + M[5] * sg[m][2]);
      f3[m] = (T)(M[6] * sg[m][0] + M[7] * sg[m][1] + M[8] * sg[m][2]);
    }
  }
  __syncthreads();  // every read of the state planes is done
  if (sact) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      lds[m * PL + t] = f1[m];
      lds[(3 + m) * PL + t] = f2[m];
      lds[(6 + m) * PL + t] = f3[m];
    }
  }
  __syncthreads();
  if (sact) {
    T Dti[N], Dtj[N], Dtk[N];
#pragma unroll
    for (int l = 0; l < N; ++l) {
      Dti[l] = Dm[l * N + i];
      Dtj[l] = Dm[l * N + j];
      Dtk[l] = Dm[l * N + k];
    }
    const bool first = gq < 0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      T o = 0;
#pragma unroll
      for (int l = 0; l < N; ++l) {
        o += Dti[l] * lds[m * PL + (l + N * j + P * k) * ST + lane];
        o += Dtj[l] * lds[(3 + m) * PL + (i + N * l + P * k) * ST + lane];
        o += Dtk[l] * lds[(6 + m) * PL + (i + N * j + P * l) * ST + lane];
      }
      const T y = yk[m];
      T r;
      if constexpr (MODE == kLeap)
        r = first ? (T(2) * u[m] - y - sk * o) : (y - sk * o);
      else
        r = first ? (beta == T(0) ? alpha * o : beta * y + alpha * o) : (y + alpha * o);
      bst<T>(r, ry, vo, (uint32_t)m * cs);
    }
  }
}

// thread-per-node tile: the largest power of two ST <= 16 with n^3 ST <= 1024
// threads and LDS = 9 n^3 ST sizeof(T) + 9 n^3 sizeof(TG) + kCoef sizeof(TG)
// <= 64 KiB
template <int N, typename T, typename TG>
struct ShotTileElNode {
  static constexpr int lds(int st) {
    return 9 * N * N * N * st * (int)sizeof(T) + 9 * N * N * N * (int)sizeof(TG) +
           kCoef * (int)sizeof(TG);
  }
  static constexpr int pick() {
    int st = 16;
    while (st > 1 && (lds(st) > 65536 || N * N * N * st > 1024)) st /= 2;
    return st;
  }
  static constexpr int value = (lds(1) <= 65536 && N * N * N <= 1024) ? pick() : 0;
};

// lanes: largest power of two <= 16 with n^2*ST <= 1024 threads and
// LDS = 6*n^3*ST*sizeof(T) + 9*n^3*sizeof(TG) + kCoef*sizeof(TG) <= 64 KiB
template <int N, typename T, typename TG>
struct ShotTileEl {
  static constexpr int lds(int st) {
    return 6 * N * N * N * st * (int)sizeof(T) + 9 * N * N * N * (int)sizeof(TG) +
           kCoef * (int)sizeof(TG);
  }
  static constexpr int pick() {
    int st = 16;
    while (st > 1 && (lds(st) > 65536 || N * N * st > 1024)) st /= 2;
    return st;
  }
  static constexpr int value = lds(1) <= 65536 ? pick() : 0;
};

template <int N, int ST, typename T, typename TG>
static int run_shots_el_st(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  const dim3 grid((unsigned)(ne * ((a.nshots + ST - 1) / ST)));
  const dim3 block(N * N * ST);
  const T *D = static_cast<const T *>(a.Dm);
  const T *U = static_cast<const T *>(a.U);
  T *Y = static_cast<T *>(a.Y);
  const T *s = static_cast<const T *>(a.sdiag);
  const TG *G = static_cast<const TG *>(a.G);
  const TG *cf = static_cast<const TG *>(a.coef);
  const TG *gw = static_cast<const TG *>(a.glxw);
  const bool tri = a.geo == kGeoTrilinear;
  if (tri && (!cf || !gw)) return -1002;
#define MJW_SHOTS_EL(MODE, TRI)                                                                  \
  hipLaunchKernelGGL((k_shots_el<N, ST, T, TG, MODE, TRI>), grid, block, 0, st, a.map, G, D,    \
                     a.lam, a.mu, a.E, a.e0, a.e1, U, Y, s, (T)a.alpha, (T)a.beta, a.N,        \
                     a.nshots, a.nxcd, cf, gw)
  if (a.mode == kLeap) {
    if (tri) MJW_SHOTS_EL(kLeap, true);
    else MJW_SHOTS_EL(kLeap, false);
  } else {
    if (tri) MJW_SHOTS_EL(kApply, true);
    else MJW_SHOTS_EL(kApply, false);
  }
#undef MJW_SHOTS_EL
  return (int)hipGetLastError();
}

template <int N, int ST, typename T, typename TG>
static int run_shots_el_node_st(const StiffArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  const dim3 grid((unsigned)(ne * ((a.nshots + ST - 1) / ST)));
  const dim3 block(N * N * N * ST);
  const T *D = static_cast<const T *>(a.Dm);
  const T *U = static_cast<const T *>(a.U);
  T *Y = static_cast<T *>(a.Y);
  const T *s = static_cast<const T *>(a.sdiag);
  const TG *G = static_cast<const TG *>(a.G);
  const TG *cf = static_cast<const TG *>(a.coef);
  const TG *gw = static_cast<const TG *>(a.glxw);
  const bool tri = a.geo == kGeoTrilinear;
  if (tri && (!cf || !gw)) return -1002;
#define MJW_SHOTS_ELN(MODE, TRI)                                                             \
  hipLaunchKernelGGL((k_shots_el_node<N, ST, T, TG, MODE, TRI>), grid, block, 0, st, a.map, \
                     G, D, a.lam, a.mu, a.E, a.e0, a.e1, U, Y, s, (T)a.alpha, (T)a.beta,    \
                     a.N, a.nshots, a.nxcd, cf, gw)
  if (a.mode == kLeap) {
    if (tri) MJW_SHOTS_ELN(kLeap, true);
    else MJW_SHOTS_ELN(kLeap, false);
  } else {
    if (tri) MJW_SHOTS_ELN(kApply, true);
    else MJW_SHOTS_ELN(kApply, false);
  }
#undef MJW_SHOTS_ELN
  return (int)hipGetLastError();
}

// tile -ST: one thread per (node, shot) with ST shot lanes
template <int N, typename T, typename TG>
static int run_shots_el_node_t(const StiffArgs &a, hipStream_t st, int want) {
  constexpr int SM = ShotTileElNode<N, T, TG>::value;
  if constexpr (SM < 1) {
    return -1005;
  } else {
    if (want > SM) return -1005;
    if constexpr (SM >= 16)
      if (want == 16) return run_shots_el_node_st<N, 16, T, TG>(a, st);
    if constexpr (SM >= 8)
      if (want == 8) return run_shots_el_node_st<N, 8, T, TG>(a, st);
    if constexpr (SM >= 4)
      if (want == 4) return run_shots_el_node_st<N, 4, T, TG>(a, st);
    if constexpr (SM >= 2)
      if (want == 2) return run_shots_el_node_st<N, 2, T, TG>(a, st);
    if (want == 1) return run_shots_el_node_st<N, 1, T, TG>(a, st);
    return -1005;
  }
}

template <int N, typename T, typename TG>
static int run_shots_el_t(const StiffArgs &a, hipStream_t st) {
  constexpr int SM = ShotTileEl<N, T, TG>::value;
  if (a.epb < 0) return run_shots_el_node_t<N, T, TG>(a, st, -a.epb);
  if constexpr (SM < 1) {
    return -1005;
  } else {
    int want = a.epb;
    if (want <= 0) {
      want = 1;
      while (want < a.nshots && want < SM) want *= 2;
    }
    if (want > SM) return -1005;
    if constexpr (SM >= 16)
      if (want == 16) return run_shots_el_st<N, 16, T, TG>(a, st);
    if constexpr (SM >= 8)
      if (want == 8) return run_shots_el_st<N, 8, T, TG>(a, st);
    if constexpr (SM >= 4)
      if (want == 4) return run_shots_el_st<N, 4, T, TG>(a, st);
    if constexpr (SM >= 2)
      if (want == 2) return run_shots_el_st<N, 2, T, TG>(a, st);
    if (want == 1) return run_shots_el_st<N, 1, T, TG>(a, st);
    return -1005;
  }
}

template <int N>
static int run_shots_el_n(const StiffArgs &a, hipStream_t st) {
  if (a.state == kF64) return run_shots_el_t<N, double, double>(a, st);
  if (a.geom == kG64) return run_shots_el_t<N, float, double>(a, st);
  return run_shots_el_t<N, float, float>(a, st);
}

int launch_stiffness_shots_elastic(const StiffArgs &a, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (a.order + 1) {
    case 2: return run_shots_el_n<2>(a, st);
    case 3: return run_shots_el_n<3>(a, st);
    case 4: return run_shots_el_n<4>(a, st);
    case 5: return run_shots_el_n<5>(a, st);
    case 6: return run_shots_el_n<6>(a, st);
    case 7: return run_shots_el_n<7>(a, st);
    case 8: return run_shots_el_n<8>(a, st);
    default: return -1000;
  }
}

template <int N>
static void shot_tiles_el_n(int state, int geom, std::vector<int> &v) {
  int sm, sn;
  if (state == kF64) {
    sm = ShotTileEl<N, double, double>::value;
    sn = ShotTileElNode<N, double, double>::value;
  } else if (geom == kG64) {
    sm = ShotTileEl<N, float, double>::value;
    sn = ShotTileElNode<N, float, double>::value;
  } else {
    sm = ShotTileEl<N, float, float>::value;
    sn = ShotTileElNode<N, float, float>::value;
  }
  for (int t = sm; t >= 1; t /= 2) v.push_back(t);
  for (int t = sn; t >= 1; t /= 2) v.push_back(-t);
}

std::vector<int> tile_choices_shots_elastic(int order, int state, int geom) {
  std::vector<int> v;
  switch (order + 1) {
    case 2: shot_tiles_el_n<2>(state, geom, v); break;
    case 3: shot_tiles_el_n<3>(state, geom, v); break;
    case 4: shot_tiles_el_n<4>(state, geom, v); break;
    case 5: shot_tiles_el_n<5>(state, geom, v); break;
    case 6: shot_tiles_el_n<6>(state, geom, v); break;
    case 7: shot_tiles_el_n<7>(state, geom, v); break;
    case 8: shot_tiles_el_n<8>(state, geom, v); break;
    default: break;
  }
  return v;
}

}  // namespace mjw
