// Operator matrices and index arithmetic of the element, predictor and
// element-face layouts, shared by the step and its adjoint.
#pragma once
#include "spd_device.hpp"

namespace spd {

// Operator matrices in the arithmetic type R of the kernels that use them;
// the quadrature weights stay FP64 (they scale the FP64 update).
template <int P, typename R = double> struct Ops {
  static constexpr int n = P + 1, nf = P + 2, T = P + 1;
  R s2f[nf][n];  // sp_to_fp
  R d2s[n][nf];  // dfp_to_sp
  R s2c[n][n];   // sp_to_cv
  R c2s[n][n];   // cv_to_sp
  R s2r[n][n];   // solution points -> regular subcell centres
  R inv[T][T];   // invader
  double w[T];   // w_tp
};

// ---------------------------------------------------------------------------
// index helpers
// ---------------------------------------------------------------------------
// Remainder of the non-negative index x by d > 0, with x replaced by the
// quotient: 32-bit arithmetic when x fits (a 64-bit division by a run-time
// divisor costs several times as much).
__device__ __forceinline__ int split_index(long long& x, int d) {
  if ((unsigned long long)x <= 0xffffffffull) {
    const unsigned u = (unsigned)x, q = u / (unsigned)d;
    x = (long long)q;
    return (int)(u - q * (unsigned)d);
  }
  const long long q = x / d;
  const int r = (int)(x - q * d);
  x = q;
  return r;
}

// Coordinates of element e; 32-bit division when the element count allows
// (a 64-bit division by a run-time divisor costs several times more).
template <int NDIM>
__device__ __forceinline__ void elem_coords(const Geom& g, long long e, int* c) {
  if (g.NE <= 0xffffffffll) {
    unsigned u = (unsigned)e;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      c[a] = (int)(u % (unsigned)g.N[a]);
      u /= (unsigned)g.N[a];
    }
    return;
  }
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    c[a] = (int)(e % g.N[a]);
    e /= g.N[a];
  }
}
template <int NDIM>
__device__ __forceinline__ long long elem_index(const Geom& g, const int* c) {
  long long e = 0;
#pragma unroll
  for (int a = NDIM - 1; a >= 0; --a) e = e * g.N[a] + c[a];
  return e;
}
// local point index of the line `lp` (local coordinates other than a) at
// position 0 along a; the stride along a is n^a.
template <int NDIM, int n>
__device__ __forceinline__ int line_base(int a, int lp) {
  int c[3] = {0, 0, 0};
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    if (b == a) continue;
    c[b] = lp % n;
    lp /= n;
  }
  int l = 0;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) l = l * n + c[b];
  return l;
}
template <int n> __device__ __forceinline__ int stride_of(int a) {
  int s = 1;
  for (int b = 0; b < a; ++b) s *= n;
  return s;
}

// Element-face arrays use a "face grid" per dimension a: coordinates q_b =
// E_b * n + l_b for b != a (transverse solution point) and q_a = face index in
// [0, N_a]; x fastest, then the variable, then the time node:
//   idx = ((t * V + v) * prod_b ext_b) + sum_b q_b * stride_b
template <int NDIM, int n>
__device__ __forceinline__ long long face_grid_base(const Geom& g, int a, const int* ce,
                                                    const int* cl, int f, long long* plane) {
  long long idx = 0, tot = 1;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) {
    const long long ext = (b == a) ? (long long)g.N[a] + 1 : (long long)g.N[b] * n;
    const long long q = (b == a) ? f : (long long)ce[b] * n + cl[b];
    idx = idx * ext + q;
    tot *= ext;
  }
  *plane = tot;
  return idx;
}

template <int NDIM, int n>
__device__ __forceinline__ long long stride_face(const Geom& g, int a) {
  long long st = 1;
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    if (b == a) break;
    st *= (long long)g.N[b] * n;
  }
  return st;
}

}  // namespace spd
