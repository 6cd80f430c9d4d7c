// First-order split acoustic system with PML layers (LEAP_FROG_PML):
//   D dU/dt = R^S V - S U + F(t) b,   B dV/dt = R^V U - S^V V,
// U = (u, u*, u**) on the nodes (split copies on PML nodes only), V at the
// quadrature points (the GL nodes) of every element.  One step:
//   tail: traces of u^n, F_n b added to P (P holds R^S V^{n+1/2});
//   k_pml_block / k_pml_scalar: U^{n+1} = (D + dt/2 S)^{-1} ((D - dt/2 S) U^n
//     + dt P) node by node (d x d blocks on PML nodes), P cleared;
//   k_pml_elem, one launch per element colour: V^{n+3/2} at each quadrature
//     point from grad u^{n+1}, then P += R^S V^{n+3/2} (scatter-free within a
//     colour).
// Arithmetic in FP64; U, V, P in the state precision T.
#include <hip/hip_runtime.h>

#include "mjw_device.hpp"
#include "mjw_internal.hpp"

namespace mjw {

template <int N, int DIM>
struct PmlShape {
  static constexpr int P = N * N;
  static constexpr int ND = DIM == 2 ? P : P * N;
  static constexpr int THREADS = ((ND + 63) / 64) * 64;
};

// one element per block, one thread per element node q = i + N j (+ N^2 k);
// point arrays are plane-major like the element map: pidx = (k E + e) N^2 + i + N j
template <int N, int DIM, typename T>
__global__ __launch_bounds__((PmlShape<N, DIM>::THREADS)) void k_pml_elem(
    const int32_t *__restrict__ map, const double *__restrict__ G, const double *__restrict__ B,
    const double *__restrict__ tau, const uint8_t *__restrict__ pml, const double *__restrict__ Dm,
    int64_t E, int64_t e0, int64_t e1, T *__restrict__ U, T *__restrict__ V, T *__restrict__ Pr,
    int64_t Nn, double dt) {
  constexpr int P = PmlShape<N, DIM>::P, ND = PmlShape<N, DIM>::ND;
  __shared__ double su[ND];
  __shared__ double fl[DIM][DIM][ND];
  const int q = threadIdx.x;
  const int64_t e = e0 + blockIdx.x;
  if (e >= e1) return;  // uniform over the block
  const bool act = q < ND;
  const int i = q % N, j = (q / N) % N, k = q / P;
  const int64_t pidx = ((int64_t)k * E + e) * P + (q % P);
  const int64_t EN = E * ND;
  const int nc = pml[e] ? DIM : 1;
  int32_t gid = 0;
  if (act) {
    gid = map[pidx] & 0x7fffffff;
    su[q] = (double)U[gid];
  }
  __syncthreads();
  if (act) {
    double gr[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < N; ++m) {
      gr[0] += Dm[i * N + m] * su[m + N * j + P * k];
      gr[1] += Dm[j * N + m] * su[i + N * m + P * k];
      if constexpr (DIM == 3) gr[2] += Dm[k * N + m] * su[i + N * j + P * m];
    }
    double g[DIM][DIM];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int c = 0; c < DIM; ++c) g[a][c] = G[(int64_t)(a * DIM + c) * EN + pidx];
    const double Bq = B[pidx];
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      // (R^V u)_a = (w J DF^{-T} grad u)_a = sum_c G[c][a] du/dxi_c
      double rv = 0.0;
#pragma unroll
      for (int c = 0; c < DIM; ++c) rv += g[c][a] * gr[c];
      const double sv = nc > 1 ? tau[(int64_t)a * EN + pidx] * Bq : 0.0;
      const double pv = dt * rv + (Bq - 0.5 * dt * sv) * (double)V[(int64_t)a * EN + pidx];
      v[a] = pv * (1.0 / (Bq + 0.5 * dt * sv));
      V[(int64_t)a * EN + pidx] = (T)v[a];
    }
    // fluxes -G P_c v: P_0 = I, P_1 = diag(1, -1[, 0]), P_2 = diag(1, 0, -1)
    for (int c = 0; c < nc; ++c) {
      double pv[3] = {v[0], v[1], v[2]};
      if (c == 1) { pv[1] = -v[1]; pv[2] = 0.0; }
      if (c == 2) { pv[1] = 0.0; pv[2] = -v[2]; }
#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        double s = 0.0;
#pragma unroll
        for (int cc = 0; cc < DIM; ++cc) s += g[a][cc] * pv[cc];
        fl[c][a][q] = -s;
      }
    }
  }
  __syncthreads();
  if (act) {
    for (int c = 0; c < nc; ++c) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < N; ++m) {
        s += Dm[m * N + i] * fl[c][0][m + N * j + P * k];
        s += Dm[m * N + j] * fl[c][1][i + N * m + P * k];
        if constexpr (DIM == 3) s += Dm[m * N + k] * fl[c][2][i + N * j + P * m];
      }
      T *pr = Pr + (int64_t)c * Nn + gid;
      *pr = (T)((double)*pr + s);
    }
  }
}

// nodes of PML elements: x = Pinv (A U + dt P) over the d copies
template <int DIM, typename T>
__global__ void k_pml_block(int64_t npn, const int32_t *__restrict__ node,
                            const double *__restrict__ A, const double *__restrict__ Pinv,
                            const uint8_t *__restrict__ dir, T *__restrict__ U, T *__restrict__ Pr,
                            int64_t Nn, double dt) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npn;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = node[t];
    double y[DIM], u[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) u[c] = (double)U[c * Nn + g];
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
      double s = dt * (double)Pr[r * Nn + g];
#pragma unroll
      for (int c = 0; c < DIM; ++c) s += A[(t * DIM + r) * DIM + c] * u[c];
      y[r] = s;
    }
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
      double x = 0.0;
#pragma unroll
      for (int c = 0; c < DIM; ++c) x += Pinv[(t * DIM + r) * DIM + c] * y[c];
      U[r * Nn + g] = (T)((r == 0 && dir[g]) ? 0.0 : x);
      Pr[r * Nn + g] = (T)0;
    }
  }
}

// the other nodes: u = inv (a u + dt P), inv = 0 on Dirichlet nodes
template <typename T>
__global__ void k_pml_scalar(int64_t Nn, const double *__restrict__ a,
                             const double *__restrict__ inv, const uint8_t *__restrict__ blk,
                             T *__restrict__ U, T *__restrict__ Pr, double dt) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < Nn;
       g += (int64_t)gridDim.x * blockDim.x) {
    if (blk[g]) continue;
    U[g] = (T)(inv[g] * (a[g] * (double)U[g] + dt * (double)Pr[g]));
    Pr[g] = (T)0;
  }
}

static unsigned grid_nodes(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

template <int N, int DIM, typename T>
static void elem_launch(const PmlArgs &a, hipStream_t st) {
  const int64_t ne = a.e1 - a.e0;
  hipLaunchKernelGGL((k_pml_elem<N, DIM, T>), dim3((unsigned)ne),
                     dim3(PmlShape<N, DIM>::THREADS), 0, st, a.map, a.G, a.B, a.tau, a.pml, a.Dm,
                     a.E, a.e0, a.e1, static_cast<T *>(a.U), static_cast<T *>(a.V),
                     static_cast<T *>(a.Pr), a.N, a.dt);
}

template <int DIM, typename T>
static int elem_dispatch(const PmlArgs &a, hipStream_t st) {
  switch (a.order) {
    case 1: elem_launch<2, DIM, T>(a, st); break;
    case 2: elem_launch<3, DIM, T>(a, st); break;
    case 3: elem_launch<4, DIM, T>(a, st); break;
    case 4: elem_launch<5, DIM, T>(a, st); break;
    case 5: elem_launch<6, DIM, T>(a, st); break;
    case 6: elem_launch<7, DIM, T>(a, st); break;
    case 7: elem_launch<8, DIM, T>(a, st); break;
    default: return -1;
  }
  return (int)hipGetLastError();
}

int launch_pml_elem(const PmlArgs &a, void *stream) {
  if (a.e1 <= a.e0) return 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a.dim == 2) return a.state == kF64 ? elem_dispatch<2, double>(a, st)
                                         : elem_dispatch<2, float>(a, st);
  return a.state == kF64 ? elem_dispatch<3, double>(a, st) : elem_dispatch<3, float>(a, st);
}

int launch_pml_nodes(const PmlArgs &a, void *stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a.npn > 0) {
    const unsigned gb = grid_nodes(a.npn);
    if (a.dim == 2) {
      if (a.state == kF64)
        hipLaunchKernelGGL((k_pml_block<2, double>), dim3(gb), dim3(256), 0, st, a.npn, a.pnode,
                           a.A, a.Pinv, a.dir, (double *)a.U, (double *)a.Pr, a.N, a.dt);
      else
        hipLaunchKernelGGL((k_pml_block<2, float>), dim3(gb), dim3(256), 0, st, a.npn, a.pnode,
                           a.A, a.Pinv, a.dir, (float *)a.U, (float *)a.Pr, a.N, a.dt);
    } else {
      if (a.state == kF64)
        hipLaunchKernelGGL((k_pml_block<3, double>), dim3(gb), dim3(256), 0, st, a.npn, a.pnode,
                           a.A, a.Pinv, a.dir, (double *)a.U, (double *)a.Pr, a.N, a.dt);
      else
        hipLaunchKernelGGL((k_pml_block<3, float>), dim3(gb), dim3(256), 0, st, a.npn, a.pnode,
                           a.A, a.Pinv, a.dir, (float *)a.U, (float *)a.Pr, a.N, a.dt);
    }
  }
  const unsigned gn = grid_nodes(a.N);
  if (a.state == kF64)
    hipLaunchKernelGGL(k_pml_scalar<double>, dim3(gn), dim3(256), 0, st, a.N, a.sa, a.sinv,
                       a.blk, (double *)a.U, (double *)a.Pr, a.dt);
  else
    hipLaunchKernelGGL(k_pml_scalar<float>, dim3(gn), dim3(256), 0, st, a.N, a.sa, a.sinv,
                       a.blk, (float *)a.U, (float *)a.Pr, a.dt);
  return (int)hipGetLastError();
}

}  // namespace mjw
