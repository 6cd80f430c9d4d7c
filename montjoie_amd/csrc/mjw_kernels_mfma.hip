// Interleaved acoustic shots (FP32 state) with the sum-factorised
// contractions on the matrix cores, gfx950.
//
// Same operator as k_shots3d (mjw_kernels_shots.hip, restating
// MltAdd_ElementH1_Scalar, Computation/ProdMatVectScalarH1.cxx:256-668): per
// element, the three xi-derivatives of u at every node and shot, the flux
// G b per node and shot, and the three transposed contractions summed into
// the update.  One block takes one element and a tile of 16 shots.
//
// Each contraction out[s][c][line] = sum_l Dx[c][l] in[s][l][line] (a line is
// the n nodes along the contracted axis) is a sequence of n rank-1 updates on
// v_mfma_f32_4x4x1f32: 16 independent 4x4 blocks per wavefront, rows = 4
// shots, columns = 4 output indices, K = the contracted index.  The 16
// blocks cover 2 lines x 4 shot quads x 2 column quads (n <= 8), so n
// instructions finish 2 lines of 16 shots.  Lane 4b + r of block b supplies
// row r (its shot's value at index l, read from LDS) and lane 4b + c supplies
// column c (a coefficient of D held in registers); lane 4b + c then holds the
// results of shots 4 rb .. 4 rb + 3 at output index c, written to LDS as one
// 16-byte quad.  The state planes are [node][16 shots].
//
// Stages: stage U (and build the geometry) | gradient contractions -> planes
// 1..3 | flux per (node, shot quad), in place | divergence contractions, in
// place (a task reads and writes only its own two lines) | update per (node,
// shot quad): o = (o_1 + o_2) + o_3.  The gradients equal the vector
// kernel's bit for bit (one fused multiply-add per K step, in order); the
// divergence sums its three directions in another order, so results agree
// with the vector kernels to FP32 rounding.
#include <hip/hip_runtime.h>

#include "mjw_device.hpp"
#include "mjw_internal.hpp"
#include "mjw_tri.hpp"

namespace mjw {

typedef float f4 __attribute__((ext_vector_type(4)));

template <int N>
struct MfmaShape {
  static constexpr int NT = N * N * N >= 125 ? 512 : 256;  // threads per block
};

template <int N, typename TG>
constexpr int mfma_lds_bytes() {
  return 4 * N * N * N * 16 * 4 + N * N * N * 10 * (int)sizeof(TG) + N * N * N * 4 +
         kCoef * (int)sizeof(TG);
}

template <int N, typename TG, int MODE, bool TRI>
__global__ __launch_bounds__(MfmaShape<N>::NT) void k_shots3d_mfma(
    const int32_t *__restrict__ map, const TG *__restrict__ G, const float *__restrict__ Dm,
    int64_t E, int64_t e0, int64_t e1, const float *__restrict__ U, float *__restrict__ Y,
    const float *__restrict__ sdiag, float alpha, float beta, int nshots, int nxcd,
    const TG *__restrict__ coef, const TG *__restrict__ glxw) {
  constexpr int P = N * N, NN = P * N, NT = MfmaShape<N>::NT, NW = NT / 64;
  constexpr int NQ = NN * 4;        // (node, shot quad) tasks
  constexpr int NPP = (P + 1) / 2;  // pairs of lines per direction
  constexpr int NG = TRI ? 10 : 6;  // per node: r0 r1 r2 f, or the 6 tensor entries
  constexpr int NUT = (NQ + NT - 1) / NT;
  using A = TG;
  __shared__ __align__(16) float pl[4][NN * 16];  // U; then per direction G -> flux -> div
  __shared__ A lg[NN * NG];
  __shared__ int32_t lmap[NN];
  __shared__ TG lc[TRI ? kCoef : 1];
  const int t = threadIdx.x, w = t / 64, lane = t % 64;
  const unsigned ntile = (unsigned)((nshots + 15) / 16);
  const unsigned lb = xcd_block(blockIdx.x, gridDim.x, nxcd);
  const int64_t e = e0 + (int64_t)(lb / ntile);
  const int s0 = (int)(lb % ntile) * 16;
  if (e >= e1) return;  // whole block (grid = elements x tiles: never taken)
  if constexpr (TRI)
    for (int c = t; c < kCoef; c += NT) lc[c] = coef[e * kCoef + c];
  for (int q = t; q < NN; q += NT) lmap[q] = map[((int64_t)(q / P) * E + e) * P + q % P];
  __syncthreads();
  // the element's geometry, one thread per node
  for (int q = t; q < NN; q += NT) {
    const int gij = q % P, gk = q / P;
    A *g = lg + q * NG;
    if constexpr (TRI) {
      const int gi = gij % N, gj = gij / N;
      TriCol<A> tc;
      tri_setup<A, TG>(lc, (A)glxw[gi], (A)glxw[gj], (A)glxw[N + gi] * (A)glxw[N + gj], true, tc);
      tri_rows_f<A>(tc, (A)glxw[gk], (A)glxw[N + gk], g, g + 3, g + 6, g[9]);
    } else {
#pragma unroll
      for (int c = 0; c < 6; ++c) g[c] = (A)G[((int64_t)(gk * 6 + c) * E + e) * P + gij];
    }
  }
  // U of the tile's shots at every node (0 beyond the last shot)
  const rsrc_t ru = mkres(U), ry = mkres(Y), rs = mkres(sdiag);
  for (int x = t; x < NN * 16; x += NT) {
    const int q = x / 16, s = s0 + x % 16;
    float v = 0.f;
    if (s < nshots)
      v = bld<float>(ru, ((uint32_t)(lmap[q] & 0x7fffffff) * (uint32_t)nshots + (uint32_t)s) * 4u,
                     0u);
    pl[0][x] = v;
  }
  // previous values and dt^2/M of this thread's update tasks, in flight
  float yq[NUT][4], sq[NUT];
#pragma unroll
  for (int it = 0; it < NUT; ++it) {
    const int x = t + it * NT;
    sq[it] = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) yq[it][r] = 0.f;
    if (x < NQ) {
      const uint32_t g = (uint32_t)(lmap[x >> 2] & 0x7fffffff);
      const int sh = s0 + 4 * (x & 3);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (sh + r < nshots)
          yq[it][r] = bld<float>(ry, (g * (uint32_t)nshots + (uint32_t)(sh + r)) * 4u, 0u);
      if constexpr (MODE == kLeap) sq[it] = bld<float>(rs, g * 4u, 0u);
    }
  }
  // this lane's block: line select, shot quad, column quad
  const int q4 = lane & 3, blk = lane >> 2;
  const int ps = blk >> 3, rb = (blk >> 1) & 3, cb = blk & 1;
  const int col = 4 * cb + q4;  // output index this lane holds
  const int shA = 4 * rb + q4;  // shot this lane supplies as a row
  float Bg[N], Bd[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    Bg[l] = col < N ? Dm[col * N + l] : 0.f;  // gradient: out[c] = sum_l D[c][l] in[l]
    Bd[l] = col < N ? Dm[l * N + col] : 0.f;  // divergence: out[c] = sum_l D[l][c] in[l]
  }
  // node of index l on line p of direction d (0: xi_1, lines (j, k); 1: xi_2,
  // lines (i, k); 2: xi_3, lines (i, j))
  auto node = [](int d, int p, int l) {
    return d == 0 ? l + N * p : (d == 1 ? (p % N) + N * l + P * (p / N) : p + P * l);
  };
  auto contract = [&](bool grad) {
    for (int task = w; task < 3 * NPP; task += NW) {  // wave-uniform
      const int d = task / NPP, p = 2 * (task % NPP) + ps;
      const bool pok = p < P;
      const float *src = grad ? pl[0] : pl[1 + d];
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int l = 0; l < N; ++l) {
        const float a = pok ? src[node(d, pok ? p : 0, l) * 16 + shA] : 0.f;
        acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a, grad ? Bg[l] : Bd[l], acc, 0, 0, 0);
      }
      if (pok && col < N)
        *reinterpret_cast<f4 *>(&pl[1 + d][node(d, p, col) * 16 + 4 * rb]) = acc;
    }
  };
  __syncthreads();  // U staged, geometry built
  contract(true);
  __syncthreads();
  // flux per (node, shot quad), in place
  for (int x = t; x < NQ; x += NT) {
    const int q = x >> 2, o = q * 16 + 4 * (x & 3);
    const A *g = lg + q * NG;
    const f4 b1 = *reinterpret_cast<const f4 *>(&pl[1][o]);
    const f4 b2 = *reinterpret_cast<const f4 *>(&pl[2][o]);
    const f4 b3 = *reinterpret_cast<const f4 *>(&pl[3][o]);
    f4 v1, v2, v3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const A c1 = b1[r], c2 = b2[r], c3 = b3[r];
      if constexpr (TRI) {
        A w1, w2, w3;
        tri_flux<A>(g, g + 3, g + 6, g[9], c1, c2, c3, w1, w2, w3);
        v1[r] = (float)w1;
        v2[r] = (float)w2;
        v3[r] = (float)w3;
      } else {
        v1[r] = (float)(g[0] * c1 + g[1] * c2 + g[2] * c3);
        v2[r] = (float)(g[1] * c1 + g[3] * c2 + g[4] * c3);
        v3[r] = (float)(g[2] * c1 + g[4] * c2 + g[5] * c3);
      }
    }
    *reinterpret_cast<f4 *>(&pl[1][o]) = v1;
    *reinterpret_cast<f4 *>(&pl[2][o]) = v2;
    *reinterpret_cast<f4 *>(&pl[3][o]) = v3;
  }
  __syncthreads();
  contract(false);
  __syncthreads();
  // update per (node, shot quad)
#pragma unroll
  for (int it = 0; it < NUT; ++it) {
    const int x = t + it * NT;
    if (x >= NQ) continue;
    const int q = x >> 2, o = q * 16 + 4 * (x & 3), sh = s0 + 4 * (x & 3);
    const int32_t m = lmap[q];
    const uint32_t g = (uint32_t)(m & 0x7fffffff);
    const bool first = m < 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (sh + r >= nshots) continue;
      const float ov = (pl[1][o + r] + pl[2][o + r]) + pl[3][o + r];
      const float y = yq[it][r];
      float res;
      if constexpr (MODE == kLeap)
        res = first ? (2.f * pl[0][o + r] - y - sq[it] * ov) : (y - sq[it] * ov);
      else
        res = first ? (beta == 0.f ? alpha * ov : beta * y + alpha * ov) : (y + alpha * ov);
      bst<float>(res, ry, (g * (uint32_t)nshots + (uint32_t)(sh + r)) * 4u, 0u);
    }
  }
}

template <int N, typename TG>
static int run_mfma(const StiffArgs &a, hipStream_t st) {
  static_assert(mfma_lds_bytes<N, TG>() <= 160 * 1024, "LDS");
  const int64_t ne = a.e1 - a.e0;
  if (ne <= 0) return 0;
  const dim3 grid((unsigned)(ne * ((a.nshots + 15) / 16)));
  const dim3 block(MfmaShape<N>::NT);
  const float *D = static_cast<const float *>(a.Dm);
  const float *U = static_cast<const float *>(a.U);
  float *Y = static_cast<float *>(a.Y);
  const float *s = static_cast<const float *>(a.sdiag);
  const TG *G = static_cast<const TG *>(a.G);
  const TG *cf = static_cast<const TG *>(a.coef);
  const TG *gw = static_cast<const TG *>(a.glxw);
  const bool tri = a.geo == kGeoTrilinear;
  if (tri && (!cf || !gw)) return -1002;
#define MJW_MFMA(MODE, TRI)                                                                  \
  hipLaunchKernelGGL((k_shots3d_mfma<N, TG, MODE, TRI>), grid, block, 0, st, a.map, G, D, a.E, \
                     a.e0, a.e1, U, Y, s, (float)a.alpha, (float)a.beta, a.nshots, a.nxcd, cf,  \
                     gw)
  if (a.mode == kLeap) {
    if (tri) MJW_MFMA(kLeap, true);
    else MJW_MFMA(kLeap, false);
  } else {
    if (tri) MJW_MFMA(kApply, true);
    else MJW_MFMA(kApply, false);
  }
#undef MJW_MFMA
  return (int)hipGetLastError();
}

template <int N>
static int run_mfma_n(const StiffArgs &a, hipStream_t st) {
  if (a.geom == kG64) return run_mfma<N, double>(a, st);
  return run_mfma<N, float>(a, st);
}

bool shots_mfma_supported(int order, int state) { return state != kF64 && order >= 1 && order <= 6; }

int launch_stiffness_shots_mfma(const StiffArgs &a, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!shots_mfma_supported(a.order, a.state)) return -1005;
  switch (a.order + 1) {
    case 2: return run_mfma_n<2>(a, st);
    case 3: return run_mfma_n<3>(a, st);
    case 4: return run_mfma_n<4>(a, st);
    case 5: return run_mfma_n<5>(a, st);
    case 6: return run_mfma_n<6>(a, st);
    case 7: return run_mfma_n<7>(a, st);
    default: return -1000;
  }
}

}  // namespace mjw
