// Per-node geometry of straight-sided hexahedra, rebuilt in the kernels from
// the trilinear map of the element's vertices.
//
// x(xi) = a0 + a1 xi1 + a2 xi2 + a3 xi3 + a4 xi1 xi2 + a5 xi2 xi3 + a6 xi1 xi3
//         + a7 xi1 xi2 xi3.  On one (i, j) column the columns of DF at layer
// zeta = x_k are c0 = p0 + zeta q0, c1 = p1 + zeta q1 and c2 (fixed), with
// p0 = a1 + a4 x_j, q0 = a6 + a7 x_j, p1 = a2 + a4 x_i, q1 = a5 + a7 x_i,
// c2 = a3 + a5 x_j + a6 x_i + a7 x_i x_j.  The rows of the adjugate
// J DF^{-1} are r0 = c1 x c2, r1 = c2 x c0, r2 = c0 x c1 and J = c0 . r0, so
// each is a polynomial in zeta whose coefficients are set once per column:
//   r0 = P0 + zeta Q0,  r1 = S0 + zeta S1,  r2 = T0 + zeta (T1 + zeta T2),
//   J  = j0 + zeta (j1 + zeta j2).
// Per node that leaves 14 FMAs for r and J instead of two cross products and
// a dot product.
#pragma once

#include "mjw_internal.hpp"

namespace mjw {

template <typename A>
struct TriCol {
  A P0[3], Q0[3], S0[3], S1[3], T0[3], T1[3], T2[3];
  A j0, j1, j2;
  A f;  // acoustic: mu w_i w_j; elastic: w_i w_j
};

template <typename A>
__device__ __forceinline__ void cross3(const A *a, const A *b, A *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename A>
__device__ __forceinline__ A dot3(const A *a, const A *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// c: the element's kCoef coefficients (a1..a7, then mu); f = c[21] * fij
// when with_mu, fij otherwise
template <typename A, typename TG>
__device__ __forceinline__ void tri_setup(const TG *c, A xi, A eta, A fij, bool with_mu,
                                          TriCol<A> &t) {
  A p0[3], q0[3], p1[3], q1[3], c2[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const A a1 = c[d], a2 = c[3 + d], a3 = c[6 + d], a4 = c[9 + d], a5 = c[12 + d],
            a6 = c[15 + d], a7 = c[18 + d];
    p0[d] = a1 + a4 * eta;
    q0[d] = a6 + a7 * eta;
    p1[d] = a2 + a4 * xi;
    q1[d] = a5 + a7 * xi;
    c2[d] = a3 + a5 * eta + a6 * xi + a7 * (xi * eta);
  }
  cross3(p1, c2, t.P0);
  cross3(q1, c2, t.Q0);
  cross3(c2, p0, t.S0);
  cross3(c2, q0, t.S1);
  cross3(p0, p1, t.T0);
  A u[3], v[3];
  cross3(p0, q1, u);
  cross3(q0, p1, v);
#pragma unroll
  for (int d = 0; d < 3; ++d) t.T1[d] = u[d] + v[d];
  cross3(q0, q1, t.T2);
  t.j0 = dot3(p0, t.P0);
  t.j1 = dot3(p0, t.Q0) + dot3(q0, t.P0);
  t.j2 = dot3(q0, t.Q0);
  t.f = with_mu ? (A)c[21] * fij : fij;
}

template <typename A>
__device__ __forceinline__ void tri_rows(const TriCol<A> &t, A zeta, A r0[3], A r1[3], A r2[3],
                                         A &J) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r0[d] = t.P0[d] + zeta * t.Q0[d];
    r1[d] = t.S0[d] + zeta * t.S1[d];
    r2[d] = t.T0[d] + zeta * (t.T1[d] + zeta * t.T2[d]);
  }
  J = t.j0 + zeta * (t.j1 + zeta * t.j2);
}

// f / J: FP32 through the hardware reciprocal (1 ulp), FP64 exact division
__device__ __forceinline__ float tri_ratio(float f, float J) {
  return f * __builtin_amdgcn_rcpf(J);
}
__device__ __forceinline__ double tri_ratio(double f, double J) { return f / J; }

// adjugate rows r0, r1, r2 and f = mu w / J at layer zeta
template <typename A>
__device__ __forceinline__ void tri_rows_f(const TriCol<A> &t, A zeta, A wk, A r0[3], A r1[3],
                                           A r2[3], A &f) {
  A J;
  tri_rows(t, zeta, r0, r1, r2, J);
  f = tri_ratio(t.f * wk, J);
}

// acoustic flux v = G b with G = mu w J DF^{-1} DF^{-T}, evaluated as
// f R (R^T b), R the matrix of rows r_a
template <typename A>
__device__ __forceinline__ void tri_flux(const A *r0, const A *r1, const A *r2, A f, A b1, A b2,
                                         A b3, A &v1, A &v2, A &v3) {
  A w[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) w[d] = b1 * r0[d] + b2 * r1[d] + b3 * r2[d];
  v1 = f * dot3(r0, w);
  v2 = f * dot3(r1, w);
  v3 = f * dot3(r2, w);
}

template <typename A>
__device__ __forceinline__ void tri_apply(const TriCol<A> &t, A zeta, A wk, A b1, A b2, A b3,
                                          A &v1, A &v2, A &v3) {
  A r0[3], r1[3], r2[3], f;
  tri_rows_f(t, zeta, wk, r0, r1, r2, f);
  tri_flux(r0, r1, r2, f, b1, b2, b3, v1, v2, v3);
}

// acoustic tensor at layer zeta (6 components 00 01 02 11 12 22)
template <typename A>
__device__ __forceinline__ void tri_tensor(const TriCol<A> &t, A zeta, A wk, A g[6]) {
  A r0[3], r1[3], r2[3], J;
  tri_rows(t, zeta, r0, r1, r2, J);
  const A f = tri_ratio(t.f * wk, J);
  g[0] = f * dot3(r0, r0);
  g[1] = f * dot3(r0, r1);
  g[2] = f * dot3(r0, r2);
  g[3] = f * dot3(r1, r1);
  g[4] = f * dot3(r1, r2);
  g[5] = f * dot3(r2, r2);
}

// elastic factor M = sqrt(J w) DF^{-1} = sqrt(w / J) adj(DF), rows r_b:
// M[b*3 + a] = sqrt(w / J) r_b[a]
template <typename A>
__device__ __forceinline__ void tri_m(const TriCol<A> &t, A zeta, A wk, A M[9]) {
  A r0[3], r1[3], r2[3], J;
  tri_rows(t, zeta, r0, r1, r2, J);
  const A f = sqrt(tri_ratio(t.f * wk, J));
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    M[a] = f * r0[a];
    M[3 + a] = f * r1[a];
    M[6 + a] = f * r2[a];
  }
}

}  // namespace mjw
