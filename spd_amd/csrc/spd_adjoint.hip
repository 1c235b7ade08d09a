// Reverse-mode derivative (adjoint) of the SD-ADER step.  The reverse sweep
// is FP64; in the FP32-bulk form the tape's predictor iterates are the
// form's FP32 values (the sweep linearises the FP64 formulas there).
//
// The step is differentiated with its trouble flags, blending coefficients
// and dt frozen (taken from a recorded forward pass): with those fixed the
// step is a composition of linear operators (Lagrange interpolation and
// differentiation, the ADER time matrix, sp <-> cv transforms, transverse
// subcell averages, the FV update and the blend) and pointwise or stencil
// maps (Euler flux, prims / cons, the element-face and fallback Riemann
// solvers, the MUSCL-Hancock state).  The linear parts are transposed by
// hand; the Jacobian-transpose products of the pointwise maps are formed
// from forward-mode dual numbers run through the same device functions the
// step uses (spd_device.hpp), one pass per input, so every branch (HLLC
// state selection, limiter case, sound-speed clamp) is the one the forward
// pass took.
//
// Forward (src/sdader_simulator.py:160-200, 288-349, src/fv_simulator.py:
// 167-205), with U_ad^(k) the k-th Picard iterate (U_ad^(0) = U_sp at every
// node), D(.) the SD spatial operator (element-face Riemann solves + flux-point
// derivative, src/sdader_simulator.py:148-158,213-240):
//   U_ad^(k)   = U_sp - dt * invader . D(U_ad^(k-1))               k = 1..m
//   FV:  U^0   = sp_to_cv(U_sp);   for substep i (dt_i = w_i dt):
//        U^{i+1} = U^i - dt_i sum_a Delta_a[th F_FB,a(U^i) + (1-th) F_HO,a(U_ad^(m), i)] / dx
//        U_sp' = cv_to_sp(U^T)
//   SD:  U_sp' = U_sp - dt sum_t w_t D(U_ad^(m))[t]
// The reverse sweep runs these equations backwards; the tape holds the
// iterates U_ad^(1..m), the FV states U^i and the coefficients th of every
// substep (recorded by spd_step, spd_step.hip).
#include "spd_ctx.hpp"

namespace spd {
namespace ad {

// Dual number: value and one directional derivative.
struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(Dual a, Dual b) { return a.v == b.v; }
__device__ __forceinline__ bool operator!=(Dual a, Dual b) { return a.v != b.v; }
__device__ __forceinline__ Dual fabs(Dual a) { return (a.v < 0.0) ? -a : a; }
__device__ __forceinline__ Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, a.d * 0.5 / s);
}

// Accumulate into device memory with the hardware FP64 atomic add: every
// buffer the adjoint adds into is hipMalloc'd (coarse-grained) device memory,
// for which the instruction is supported (not fine-grained host memory).
__device__ __forceinline__ void acc_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// xbar[k] = sum_j ybar[j] * d f_j / d x_k at x: one forward-mode pass per input.
template <int NI, int NO, typename F>
__device__ __forceinline__ void vjp(const double* x, const double* ybar, double* xbar, F&& f) {
#pragma unroll 1
  for (int k = 0; k < NI; ++k) {
    Dual xd[NI], yd[NO];
#pragma unroll
    for (int i = 0; i < NI; ++i) xd[i] = Dual(x[i], i == k ? 1.0 : 0.0);
    f(xd, yd);
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < NO; ++j) acc += ybar[j] * yd[j].d;
    // a store at a constant index: xbar (the caller's array) stays in
    // registers instead of going to scratch memory for a run-time index
#pragma unroll
    for (int i = 0; i < NI; ++i)
      if (i == k) xbar[i] = acc;
  }
}

}  // namespace ad

using ad::Dual;

// Dynamic LDS of the reverse kernels (typed at each use).
extern __shared__ __attribute__((aligned(16))) unsigned char spd_adj_lds[];

// ---------------------------------------------------------------------------
// Transpose of the tensor-product transform with the n x n matrix M (which: 0
// cv_to_sp, 1 sp_to_cv) on every element: out[qo] = sum_qi prod_b M[qi_b][qo_b]
// in[qi]; in_fv / out_fv select the FV layout (src/transforms.py:27-45).
// ---------------------------------------------------------------------------
template <int NDIM, int P>
__global__ void __launch_bounds__(256)
k_adj_tensor(Ops<P> op, Geom g, int which, const double* __restrict__ in, int in_fv,
             double* __restrict__ out, int out_fv) {
  constexpr int V = Dims<NDIM>::V, n = P + 1;
  constexpr int NL = IPow<n, NDIM>::v;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.NE * V * NL) return;
  const int lo = (int)(tid % NL);
  const long long r = tid / NL;
  const long long e = r % g.NE;
  const int v = (int)(r / g.NE);
  const double(*M)[n] = which ? op.s2c : op.c2s;
  int ce[3] = {0, 0, 0};
  elem_coords<NDIM>(g, e, ce);
  auto index = [&](int fv, int l) -> long long {
    if (!fv) return ((long long)v * g.NE + e) * NL + l;
    int cc[3] = {0, 0, 0};
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      cc[b] = ce[b] * n + l % n;
      l /= n;
    }
    return (long long)v * g.Nc + cell_index<NDIM>(g, cc);
  };
  int qo[3] = {0, 0, 0};
  {
    int x = lo;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      qo[b] = x % n;
      x /= n;
    }
  }
  double acc = 0.0;
  for (int li = 0; li < NL; ++li) {
    int x = li;
    double w = 1.0;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      w *= M[x % n][qo[b]];
      x /= n;
    }
    acc += w * in[index(in_fv, li)];
  }
  out[index(out_fv, lo)] = acc;
}

// ---------------------------------------------------------------------------
// The same transposes in 3-D through LDS: a block owns ~256 points of
// consecutive elements, loads all variables in the input layout, applies M^T
// along x, then y, then z (one thread per point; the separable form of the
// single sum above, equal to it up to round-off) and writes the other layout.
// IN_FV: input in FV layout, output in element layout (else the reverse).
// ---------------------------------------------------------------------------
template <int P, int IN_FV>
__global__ void __launch_bounds__(256)
k_adj_tensor3d(Ops<P> op, Geom g, int which, const double* __restrict__ in,
               double* __restrict__ out) {
  constexpr int V = 5, n = P + 1, NL = n * n * n;
  constexpr int EB = 256 / NL > 0 ? 256 / NL : 1, NP = EB * NL;
  __shared__ double sa[V][NP], sb[V][NP];
  const long long e0 = (long long)blockIdx.x * EB;
  const int ne = (g.NE - e0 < EB) ? (int)(g.NE - e0) : EB;
  auto cell_of = [&](int el, int l) -> long long {
    int ce[3];
    elem_coords<3>(g, e0 + el, ce);
    const int cc[3] = {ce[0] * n + l % n, ce[1] * n + (l / n) % n, ce[2] * n + l / (n * n)};
    return cell_index<3>(g, cc);
  };
  for (int k = threadIdx.x; k < V * ne * NL; k += blockDim.x) {
    const int v = k / (ne * NL), r = k % (ne * NL);
    sa[v][r] = IN_FV ? in[(long long)v * g.Nc + cell_of(r / NL, r % NL)]
                     : in[((long long)v * g.NE + e0) * NL + r];
  }
  __syncthreads();
  const double(*M)[n] = which ? op.s2c : op.c2s;
  const int k = threadIdx.x;
  const bool act = k < ne * NL;
  const int el = k / NL, l = k % NL;
  const int i = l % n, j = (l / n) % n, z = l / (n * n);
  const int b0 = el * NL;
  if (act) {  // x
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[q][i] * sa[v][b0 + q + n * (j + n * z)];
      sb[v][k] = acc;
    }
  }
  __syncthreads();
  if (act) {  // y
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[q][j] * sb[v][b0 + i + n * (q + n * z)];
      sa[v][k] = acc;
    }
  }
  __syncthreads();
  if (act) {  // z, written out
    const long long ci = IN_FV ? 0 : cell_of(el, l);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[q][z] * sa[v][b0 + i + n * (j + n * q)];
      if (IN_FV)
        out[((long long)v * g.NE + e0) * NL + k] = acc;
      else
        out[(long long)v * g.Nc + ci] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Hancock states and half slopes of the cells the fallback needs (the cells
// next to a face with a non-zero coefficient), as in the step; layout
// [Mp (V), S_0 (V), ..., S_{d-1} (V)] per cell of the (Ns+2)^d grid.
// ---------------------------------------------------------------------------
template <int NDIM>
__device__ __forceinline__ double face_theta(const Geom& g, const Params& prm,
                                             const double* __restrict__ theta, int a,
                                             const int* f) {
  if (prm.godunov) return 1.0;
  long long gl = 0, gr = 0;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) {
    gl = gl * g.Ng[b] + f[b] + NGH - (b == a ? 1 : 0);
    gr = gr * g.Ng[b] + f[b] + NGH;
  }
  return nmax(theta[gl], theta[gr]);
}

// Cell c of the one-ghost grid borders a face of dimension a with a non-zero
// coefficient (or every face with godunov).
template <int NDIM>
__device__ __forceinline__ bool hs_needed(const Geom& g, const Params& prm,
                                          const double* __restrict__ theta, const int* c) {
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    bool trans_ok = true;
#pragma unroll
    for (int b = 0; b < NDIM; ++b)
      if (b != a && (c[b] < 0 || c[b] >= g.Ns[b])) trans_ok = false;
    if (!trans_ok) continue;
    int f[3] = {c[0], c[1], c[2]};
    if (c[a] >= 0 && c[a] <= g.Ns[a] && !(face_theta<NDIM>(g, prm, theta, a, f) == 0.0))
      return true;
    f[a] = c[a] + 1;
    if (f[a] <= g.Ns[a] && !(face_theta<NDIM>(g, prm, theta, a, f) == 0.0)) return true;
  }
  return false;
}

template <int NDIM>
__device__ __forceinline__ void load_stencil(const double* __restrict__ U, const Geom& g,
                                             double gamma, const int* c, double* x) {
  constexpr int V = Dims<NDIM>::V;
  load_W<NDIM>(U, g, c, x, gamma);
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    int cm[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]};
    cm[b] -= 1;
    cp[b] += 1;
    load_W<NDIM>(U, g, cm, x + (1 + b) * V, gamma);
    load_W<NDIM>(U, g, cp, x + (1 + NDIM + b) * V, gamma);
  }
}

template <int NDIM>
__global__ void __launch_bounds__(256)
k_adj_hs(Geom g, Params prm, const double* __restrict__ U, const double* __restrict__ theta,
         const double* __restrict__ dt_dev, double w_i, double* __restrict__ HS) {
  constexpr int V = Dims<NDIM>::V;
  long long n1 = 1;
  int e1[3] = {1, 1, 1};
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    e1[a] = g.Ns[a] + 2;
    n1 *= e1[a];
  }
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n1) return;
  int c[3] = {0, 0, 0};
  {
    long long x = tid;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      c[a] = (int)(x % e1[a]) - 1;
      x /= e1[a];
    }
  }
  if (!hs_needed<NDIM>(g, prm, theta, c)) return;
  double x[(1 + 2 * NDIM) * V];
  load_stencil<NDIM>(U, g, prm.gamma, c, x);
  double Mp[V], S[NDIM][V];
  hancock_state<NDIM, double>(g, prm, c, x, (const double(*)[V])(x + V),
                              (const double(*)[V])(x + (1 + NDIM) * V), (*dt_dev) * w_i, Mp, S);
#pragma unroll
  for (int v = 0; v < V; ++v) HS[(long long)v * n1 + tid] = Mp[v];
#pragma unroll
  for (int b = 0; b < NDIM; ++b)
#pragma unroll
    for (int v = 0; v < V; ++v) HS[(long long)((1 + b) * V + v) * n1 + tid] = S[b][v];
}

// ---------------------------------------------------------------------------
// Substep faces of dimension A: the flux cotangent from the substep's output
// cotangent Ub1 (transpose of U^{i+1} = U^i - dt_i sum_a Delta_a F / dx), split
// into the high-order part (1 - th) and, on faces with th != 0, the
// cotangents of the fallback face states (WL, WR) = (Mp_L + S_L, Mp_R - S_R).
// ---------------------------------------------------------------------------
template <int NDIM, int RS, int A>
__global__ void __launch_bounds__(256)
k_adj_fvface(Geom g, Params prm, int fb, const double* __restrict__ Ub1,
             const double* __restrict__ theta, const double* __restrict__ HS,
             const double* __restrict__ dt_dev, double w_i, double* __restrict__ Fhob,
             double* __restrict__ WbL, double* __restrict__ WbR) {
  constexpr int a = A;
  constexpr int V = Dims<NDIM>::V;
  long long ext[3] = {1, 1, 1}, nface = 1;
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    ext[b] = g.Ns[b] + (b == a ? 1 : 0);
    nface *= ext[b];
  }
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nface) return;
  int c[3] = {0, 0, 0};
  {
    long long x = tid;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      c[b] = split_index(x, ext[b]);
    }
  }
  const double dti = (*dt_dev) * w_i;
  double Fb[V];
#pragma unroll
  for (int v = 0; v < V; ++v) Fb[v] = 0.0;
  if (c[a] < g.Ns[a]) {
    const long long ci = cell_index<NDIM>(g, c);
    const double r = dti * g.rhfp[a][c[a] + NGH];
#pragma unroll
    for (int v = 0; v < V; ++v) Fb[v] += Ub1[v * g.Nc + ci] * r;
  }
  if (c[a] > 0) {
    int cl[3] = {c[0], c[1], c[2]};
    cl[a] -= 1;
    const long long ci = cell_index<NDIM>(g, cl);
    const double r = dti * g.rhfp[a][c[a] - 1 + NGH];
#pragma unroll
    for (int v = 0; v < V; ++v) Fb[v] -= Ub1[v * g.Nc + ci] * r;
  }
  const double th = fb ? face_theta<NDIM>(g, prm, theta, a, c) : 0.0;
#pragma unroll
  for (int v = 0; v < V; ++v) Fhob[v * nface + tid] = (th == 0.0) ? Fb[v] : (1.0 - th) * Fb[v];
  if (th == 0.0) return;
  long long hl = 0, hr = 0, n1 = 1;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) {
    hl = hl * (g.Ns[b] + 2) + c[b] + 1 - (b == a ? 1 : 0);
    hr = hr * (g.Ns[b] + 2) + c[b] + 1;
  }
#pragma unroll
  for (int b = 0; b < NDIM; ++b) n1 *= g.Ns[b] + 2;
  double x[2 * V], yb[V], xb[2 * V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    x[v] = HS[(long long)v * n1 + hl] + HS[(long long)((1 + a) * V + v) * n1 + hl];
    x[V + v] = HS[(long long)v * n1 + hr] - HS[(long long)((1 + a) * V + v) * n1 + hr];
    yb[v] = th * Fb[v];
  }
  const Dual gam(prm.gamma), mc2(prm.min_c2);
  const int lm = prm.lowmach_fv;
  ad::vjp<2 * V, V>(x, yb, xb, [&](const Dual* xd, Dual* yd) {
    Dual UL[V], UR[V];
    cons<NDIM, Dual>(xd, UL, a, gam);
    cons<NDIM, Dual>(xd + V, UR, a, gam);
    riemann<NDIM, RS, Dual>(xd, xd + V, UL, UR, yd, a, gam, mc2, lm);
  });
#pragma unroll
  for (int v = 0; v < V; ++v) {
    WbL[v * nface + tid] = xb[v];
    WbR[v * nface + tid] = xb[V + v];
  }
}

// ---------------------------------------------------------------------------
// Fallback states -> FV state: per cell of the one-ghost grid, the cotangent
// of its Hancock state (gathered from its faces), through the MUSCL-Hancock
// map of its (1 + 2d)-cell stencil, the ghost maps (periodic wrap,
// non-mirrored copies, reflective sign flips) and prims, accumulated into Ub0.
// ---------------------------------------------------------------------------
template <int NDIM>
__global__ void __launch_bounds__(256)
k_adj_hancock(Geom g, Params prm, const double* __restrict__ U, const double* __restrict__ theta,
              const double* __restrict__ dt_dev, double w_i, const double* __restrict__ WbL0,
              const double* __restrict__ WbL1, const double* __restrict__ WbL2,
              const double* __restrict__ WbR0, const double* __restrict__ WbR1,
              const double* __restrict__ WbR2, double* __restrict__ Ub0) {
  constexpr int V = Dims<NDIM>::V;
  constexpr int NI = (1 + 2 * NDIM) * V, NO = (1 + NDIM) * V;
  long long n1 = 1;
  int e1[3] = {1, 1, 1};
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    e1[a] = g.Ns[a] + 2;
    n1 *= e1[a];
  }
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n1) return;
  int c[3] = {0, 0, 0};
  {
    long long x = tid;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      c[a] = (int)(x % e1[a]) - 1;
      x /= e1[a];
    }
  }
  const double* WbL[3] = {WbL0, WbL1, WbL2};
  const double* WbR[3] = {WbR0, WbR1, WbR2};
  double yb[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) yb[k] = 0.0;
  bool any = false;
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    bool trans_ok = true;
#pragma unroll
    for (int b = 0; b < NDIM; ++b)
      if (b != a && (c[b] < 0 || c[b] >= g.Ns[b])) trans_ok = false;
    if (!trans_ok) continue;
    long long nface = 1, ext[3] = {1, 1, 1};
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      ext[b] = g.Ns[b] + (b == a ? 1 : 0);
      nface *= ext[b];
    }
    auto face_idx = [&](const int* f) {
      long long idx = 0;
#pragma unroll
      for (int b = NDIM - 1; b >= 0; --b) idx = idx * ext[b] + f[b];
      return idx;
    };
    int f[3] = {c[0], c[1], c[2]};
    // the face on the left of c: c is its right state, Mp - S_a
    if (c[a] >= 0 && c[a] <= g.Ns[a] && !(face_theta<NDIM>(g, prm, theta, a, f) == 0.0)) {
      const long long fi = face_idx(f);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const double w = WbR[a][v * nface + fi];
        yb[v] += w;
        yb[(1 + a) * V + v] -= w;
      }
      any = true;
    }
    // the face on the right of c: c is its left state, Mp + S_a
    f[a] = c[a] + 1;
    if (f[a] <= g.Ns[a] && !(face_theta<NDIM>(g, prm, theta, a, f) == 0.0)) {
      const long long fi = face_idx(f);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const double w = WbL[a][v * nface + fi];
        yb[v] += w;
        yb[(1 + a) * V + v] += w;
      }
      any = true;
    }
  }
  if (!any) return;
  double x[NI], xb[NI];
  load_stencil<NDIM>(U, g, prm.gamma, c, x);
  // The map factors as slopes dM[b][v] = slope(W[v], Wm[b][v], Wp[b][v]),
  // S[b] = dM[b] hf_b / 2, Mp = predict(W, dM): the predictor's Jacobian
  // from (1 + d) V dual passes, each slope's from three scalar passes.
  const Dual dti((*dt_dev) * w_i);
  double pin[NO], pb[NO];
#pragma unroll
  for (int v = 0; v < V; ++v) pin[v] = x[v];
#pragma unroll
  for (int b = 0; b < NDIM; ++b)
#pragma unroll
    for (int v = 0; v < V; ++v)
      pin[(1 + b) * V + v] = hancock_slope<double>(g, prm, b, c[b], x[v], x[(1 + b) * V + v],
                                                   x[(1 + NDIM + b) * V + v]);
  ad::vjp<NO, V>(pin, yb, pb, [&](const Dual* xd, Dual* yd) {
    hancock_predict<NDIM, Dual>(prm, xd, (const Dual(*)[V])(xd + V), dti, yd);
  });
#pragma unroll
  for (int v = 0; v < V; ++v) xb[v] = pb[v];
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    const double hf = g.hfp[b][c[b] + NGH];
    const int cb = c[b];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const double mb = pb[(1 + b) * V + v] + 0.5 * hf * yb[(1 + b) * V + v];
      const double s3[3] = {x[v], x[(1 + b) * V + v], x[(1 + NDIM + b) * V + v]};
      double g3[3];
      ad::vjp<3, 1>(s3, &mb, g3, [&](const Dual* xd, Dual* yd) {
        yd[0] = hancock_slope<Dual>(g, prm, b, cb, xd[0], xd[1], xd[2]);
      });
      xb[v] += g3[0];
      xb[(1 + b) * V + v] = g3[1];
      xb[(1 + NDIM + b) * V + v] = g3[2];
    }
  }
  // W of each stencil cell = flip(prims(U[map(cell)])): transpose into Ub0
  const Dual gam(prm.gamma);
  // unrolled: xb is read at k * V + v, which a run-time k would index through
  // scratch memory
#pragma unroll
  for (int k = 0; k < 1 + 2 * NDIM; ++k) {
    int cc[3] = {c[0], c[1], c[2]};
    if (k >= 1 && k <= NDIM) cc[k - 1] -= 1;
    if (k > NDIM) cc[k - 1 - NDIM] += 1;
    int mm[3];
    bool flip[3] = {false, false, false};
#pragma unroll
    for (int a = 0; a < NDIM; ++a) mm[a] = ghost_map(cc[a], g.Ns[a], g.bc[a], &flip[a]);
    const long long idx = cell_index<NDIM>(g, mm);
    double wb[V], u[V], ub[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      wb[v] = xb[k * V + v];
      u[v] = U[v * g.Nc + idx];
    }
#pragma unroll
    for (int a = 0; a < NDIM; ++a)
      if (flip[a]) wb[a + 1] = -wb[a + 1];
    ad::vjp<V, V>(u, wb, ub, [&](const Dual* ud, Dual* wd) { prims<NDIM, Dual>(ud, wd, 0, gam); });
#pragma unroll
    for (int v = 0; v < V; ++v) ad::acc_add(&Ub0[v * g.Nc + idx], ub[v]);
  }
}

// ---------------------------------------------------------------------------
// High-order FV fluxes of node t -> predictor: transpose of the transverse
// subcell averages and of the flux-point evaluation (k_hoflux / k_ho2d), all
// dimensions in one launch.  A block owns EPB elements and accumulates their
// iterate cotangent in an LDS slab; one work item per (element, line along a)
// walks the line's flux points: interior points through the Euler flux (one
// dual pass per input) into the line's solution points, end points into the
// element-face cotangent Ffb (every element face of the node is written
// exactly once).  Dimensions are processed one after the other, so the items
// of one dimension own disjoint points: no atomics.  The node's slice of
// Uadb is written once, at the end (nothing else writes it in the substep
// sweep).
// ---------------------------------------------------------------------------
template <int NDIM, int P, typename S>
__global__ void __launch_bounds__(256)
k_adj_hoflux(Ops<P> op, Geom g, Params prm, int t, const S* __restrict__ Uad,
             const double* __restrict__ Fhob0, const double* __restrict__ Fhob1,
             const double* __restrict__ Fhob2, double* __restrict__ Uadb,
             double* __restrict__ Ffb0, double* __restrict__ Ffb1, double* __restrict__ Ffb2,
             int EPB) {
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;
  double* slab = reinterpret_cast<double*>(spd_adj_lds);  // [EPB][V][NL]
  const long long e0 = (long long)blockIdx.x * EPB;
  const int ne = (int)((g.NE - e0 < EPB) ? (g.NE - e0) : EPB);
  for (int k = threadIdx.x; k < ne * V * NL; k += blockDim.x) slab[k] = 0.0;
  __syncthreads();
  const double* Fhob[3] = {Fhob0, Fhob1, Fhob2};
  double* Ffb[3] = {Ffb0, Ffb1, Ffb2};
  const Dual gam(prm.gamma);
  // unrolled: with a known at compile time the dual-number arrays of the
  // flux VJP stay in registers (a run-time a indexes them through scratch)
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    long long ext[3] = {1, 1, 1}, nface = 1;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      ext[b] = g.Ns[b] + (b == a ? 1 : 0);
      nface *= ext[b];
    }
    const int sa = stride_of<n>(a);
    for (int item = threadIdx.x; item < ne * NLP; item += blockDim.x) {
      const int el = item / NLP, lp = item % NLP;
      const long long e = e0 + el;
      int c[3] = {0, 0, 0};
      elem_coords<NDIM>(g, e, c);
      int cl[3] = {0, 0, 0};
      {
        int r = lp;
#pragma unroll
        for (int b = 0; b < NDIM; ++b) {
          if (b == a) continue;
          cl[b] = r % n;
          r /= n;
        }
      }
      const int lb = line_base<NDIM, n>(a, lp);
      double u[V][n], ub[V][n];
#pragma unroll
      for (int v = 0; v < V; ++v)
#pragma unroll
        for (int s2 = 0; s2 < n; ++s2) {
          u[v][s2] = (double)Uad[((long long)(t * V + v) * g.NE + e) * NL + lb + s2 * sa];
          ub[v][s2] = 0.0;
        }
#pragma unroll 1
      for (int f = 0; f < nf; ++f) {
        if (f == nf - 1 && c[a] != g.N[a] - 1) continue;
        // cotangent of the flux at (f, line lp): transverse-average transpose,
        // sum over the transverse subcells q of prod_b s2c[q_b][lp_b] Fhob[q]
        double Fl[V];
#pragma unroll
        for (int v = 0; v < V; ++v) Fl[v] = 0.0;
        for (int q = 0; q < NLP; ++q) {
          int qq[3] = {0, 0, 0};
          double w = 1.0;
          int r = q;
#pragma unroll
          for (int b = 0; b < NDIM; ++b) {
            if (b == a) continue;
            qq[b] = r % n;
            r /= n;
            w *= op.s2c[qq[b]][cl[b]];
          }
          long long idx = 0;
#pragma unroll
          for (int b = NDIM - 1; b >= 0; --b) {
            const long long cb = (b == a) ? (long long)c[a] * n + f : (long long)c[b] * n + qq[b];
            idx = idx * ext[b] + cb;
          }
#pragma unroll
          for (int v = 0; v < V; ++v) Fl[v] += w * Fhob[a][v * nface + idx];
        }
        if (f == 0 || f == nf - 1) {
          long long plane;
          const long long i0 = face_grid_base<NDIM, n>(g, a, c, cl, c[a] + (f == 0 ? 0 : 1),
                                                       &plane);
#pragma unroll
          for (int v = 0; v < V; ++v) Ffb[a][(long long)(t * V + v) * plane + i0] = Fl[v];
        } else {
          double M[V], Mb[V];
#pragma unroll
          for (int v = 0; v < V; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int s2 = 0; s2 < n; ++s2) acc += op.s2f[f][s2] * u[v][s2];
            M[v] = acc;
          }
          ad::vjp<V, V>(M, Fl, Mb, [&](const Dual* xd, Dual* yd) {
            Dual W[V];
            prims<NDIM, Dual>(xd, W, a, gam);
            flux<NDIM, Dual>(W, yd, a, gam);
          });
#pragma unroll
          for (int v = 0; v < V; ++v)
#pragma unroll
            for (int s2 = 0; s2 < n; ++s2) ub[v][s2] += op.s2f[f][s2] * Mb[v];
        }
      }
      double* sl = slab + (el * V) * NL + lb;
#pragma unroll
      for (int v = 0; v < V; ++v)
#pragma unroll
        for (int s2 = 0; s2 < n; ++s2) sl[v * NL + s2 * sa] += ub[v][s2];
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < ne * V * NL; k += blockDim.x) {
    const int el = k / (V * NL), v = (k / NL) % V, l = k % NL;
    Uadb[((long long)(t * V + v) * g.NE + e0 + el) * NL + l] = slab[k];
  }
}

// ---------------------------------------------------------------------------
// Element faces of dimension A (k_face_sd), first half: the flux cotangent
// through the Riemann solver and prims to the face's two traces; a boundary
// trace copied from the element's own trace (outflow, reflective) hands its
// cotangent (flipped back) to that trace.  The left trace's cotangent
// replaces Ffb in place (each thread reads and writes its own face), the
// right trace's goes to Qr; k_adj_trace gathers them into the lines.
// ---------------------------------------------------------------------------
template <int NDIM, int P, int RS, int A, typename S>
__global__ void __launch_bounds__(256)
k_adj_face(Ops<P> op, Geom g, Params prm, const S* __restrict__ src, int Tsrc,
           double* __restrict__ Ffb, double* __restrict__ Qr) {
  constexpr int a = A;
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v;
  const int Na = g.N[a];
  long long ext[3] = {1, 1, 1}, plane = 1;
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    ext[b] = (b == a) ? (long long)Na + 1 : (long long)g.N[b] * n;
    plane *= ext[b];
  }
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= plane * Tsrc) return;
  const int t = (int)(tid / plane);
  // thread -> face: the transverse coordinates vary fastest and the face
  // index along A slowest, so that adjacent lanes read adjacent lines of one
  // element; pidx is the face's index in the face grid (x fastest)
  int ce[3] = {0, 0, 0}, cl[3] = {0, 0, 0};
  int f = 0;
  long long pidx = 0;
  {
    long long x = tid % plane;
    int q[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < NDIM; ++k) {
      const int b = (k < NDIM - 1) ? (k < a ? k : k + 1) : a;
      q[b] = split_index(x, ext[b]);
    }
#pragma unroll
    for (int b = NDIM - 1; b >= 0; --b) pidx = pidx * ext[b] + q[b];
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      if (b == a) {
        f = q[b];
      } else {
        ce[b] = q[b] / n;
        cl[b] = q[b] % n;
      }
    }
  }
  int lb = 0;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) lb = lb * n + ((b == a) ? 0 : cl[b]);
  const int sa = stride_of<n>(a);
  const long long tbase = (long long)t * V * g.NE * NL;
  auto elem = [&](int ea) {
    int cc[3] = {ce[0], ce[1], ce[2]};
    cc[a] = ea;
    return elem_index<NDIM>(g, cc);
  };
  auto trace = [&](int ea, int fp, double* q) {
    const long long e = elem(ea);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const S* u = src + tbase + ((long long)v * g.NE + e) * NL + lb;
      double acc = 0.0;
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * (double)u[s * sa];
      q[v] = acc;
    }
  };
  const int bl = g.bc[a][0], br = g.bc[a][1];
  double x[2 * V];
  double* qL = x;
  double* qR = x + V;
  if (f > 0) trace(f - 1, nf - 1, qL);
  if (f < Na) trace(f, 0, qR);
  if (f == 0) {
    if (bl == BC_PERIODIC) {
      trace(Na - 1, nf - 1, qL);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) qL[v] = qR[v];
      if (bl == BC_REFLECTIVE) qL[a + 1] = -qL[a + 1];
    }
  }
  if (f == Na) {
    if (br == BC_PERIODIC) {
      trace(0, 0, qR);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) qR[v] = qL[v];
      if (br == BC_REFLECTIVE) qR[a + 1] = -qR[a + 1];
    }
  }
  double yb[V], xb[2 * V];
#pragma unroll
  for (int v = 0; v < V; ++v) yb[v] = Ffb[(long long)(t * V + v) * plane + pidx];
  const Dual gam(prm.gamma), mc2(prm.min_c2);
  const int lm = prm.lowmach_sd;
  ad::vjp<2 * V, V>(x, yb, xb, [&](const Dual* xd, Dual* yd) {
    Dual WL[V], WR[V];
    prims<NDIM, Dual>(xd, WL, a, gam);
    prims<NDIM, Dual>(xd + V, WR, a, gam);
    riemann<NDIM, RS, Dual>(WL, WR, xd, xd + V, yd, a, gam, mc2, lm);
  });
  double* qLb = xb;
  double* qRb = xb + V;
  // a copied trace hands its cotangent (flipped back) to the trace it copies
  if (f == 0 && bl != BC_PERIODIC) {
    if (bl == BC_REFLECTIVE) qLb[a + 1] = -qLb[a + 1];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      qRb[v] += qLb[v];
      qLb[v] = 0.0;
    }
  }
  if (f == Na && br != BC_PERIODIC) {
    if (br == BC_REFLECTIVE) qRb[a + 1] = -qRb[a + 1];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      qLb[v] += qRb[v];
      qRb[v] = 0.0;
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const long long k = (long long)(t * V + v) * plane + pidx;
    Ffb[k] = qLb[v];
    Qr[k] = qRb[v];
  }
}

// Second half: per (node, element, line along A), the cotangents of the
// line's two traces -- the left state of the face above it and the right
// state of the face below it, plus on periodic sides the wrapped copies at
// the domain's first / last face -- through the sp -> fp interpolation into
// the source cotangent Sb.  Each work item owns its line's points; launches
// on one stream are ordered, so the update needs no atomics.
template <int NDIM, int P, int A>
__global__ void __launch_bounds__(256)
k_adj_trace(Ops<P> op, Geom g, int Tsrc, const double* __restrict__ QlF,
            const double* __restrict__ Qr, double* __restrict__ Sb) {
  constexpr int a = A;
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)Tsrc * g.NE * NLP) return;
  const int lp = (int)(tid % NLP);
  const long long e = (tid / NLP) % g.NE;
  const int t = (int)(tid / ((long long)NLP * g.NE));
  int c[3] = {0, 0, 0};
  elem_coords<NDIM>(g, e, c);
  int cl[3] = {0, 0, 0};
  {
    int r = lp;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      if (b == a) continue;
      cl[b] = r % n;
      r /= n;
    }
  }
  const int Na = g.N[a];
  long long plane;
  const long long i0 = face_grid_base<NDIM, n>(g, a, c, cl, c[a], &plane);
  const long long sfa = stride_face<NDIM, n>(g, a);
  // a periodic first face takes its left trace from the last element, a
  // periodic last face its right trace from the first
  const bool wrapL = g.bc[a][0] == BC_PERIODIC && c[a] == Na - 1;
  const bool wrapR = g.bc[a][1] == BC_PERIODIC && c[a] == 0;
  double rb[V], lb_[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const long long k = (long long)(t * V + v) * plane + i0;
    rb[v] = QlF[k + sfa];   // face c+1: this element is its left state
    lb_[v] = Qr[k];         // face c: this element is its right state
    if (wrapL) rb[v] += QlF[k + sfa - (long long)Na * sfa];  // face 0
    if (wrapR) lb_[v] += Qr[k + (long long)Na * sfa];         // face Na
  }
  const int lbase = line_base<NDIM, n>(a, lp);
  const int sa = stride_of<n>(a);
#pragma unroll
  for (int v = 0; v < V; ++v) {
    double* ub = Sb + (long long)t * V * g.NE * NL + ((long long)v * g.NE + e) * NL + lbase;
#pragma unroll
    for (int s2 = 0; s2 < n; ++s2)
      ub[s2 * sa] += op.s2f[nf - 1][s2] * rb[v] + op.s2f[0][s2] * lb_[v];
  }
}

// ---------------------------------------------------------------------------
// SD spatial operator along dimension A (k_pred's derivative phase): the
// derivative cotangent Db -> flux-point cotangents; interior points through
// the Euler flux into Sb, end points into the element-face cotangent Ffb.
// One thread per (node, element, line along A).
// ---------------------------------------------------------------------------
template <int NDIM, int P, int A, typename S>
__global__ void __launch_bounds__(256)
k_adj_pred(Ops<P> op, Geom g, Params prm, const S* __restrict__ src, int Tsrc,
           const double* __restrict__ Db, double* __restrict__ Sb, double* __restrict__ Ffb) {
  constexpr int a = A;
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)Tsrc * g.NE * NLP) return;
  const int lp = (int)(tid % NLP);
  const long long e = (tid / NLP) % g.NE;
  const int t = (int)(tid / ((long long)NLP * g.NE));
  int c[3] = {0, 0, 0};
  elem_coords<NDIM>(g, e, c);
  const int lb = line_base<NDIM, n>(a, lp);
  const int sa = stride_of<n>(a);
  const long long tbase = (long long)t * V * g.NE * NL;
  double db[V][n];
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int s = 0; s < n; ++s) db[v][s] = Db[tbase + ((long long)v * g.NE + e) * NL + lb + s * sa];
  const double rha = g.rh[a];
  double Fb[nf][V];
#pragma unroll
  for (int fp = 0; fp < nf; ++fp)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.d2s[s][fp] * db[v][s];
      Fb[fp][v] = acc * rha;
    }
  {
    int cl[3] = {0, 0, 0};
    int r = lp;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      if (b == a) continue;
      cl[b] = r % n;
      r /= n;
    }
    long long plane;
    const long long i0 = face_grid_base<NDIM, n>(g, a, c, cl, c[a], &plane);
    const long long sfa = stride_face<NDIM, n>(g, a);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double* fb = Ffb + (long long)(t * V + v) * plane + i0;
      ad::acc_add(&fb[0], Fb[0][v]);
      ad::acc_add(&fb[sfa], Fb[nf - 1][v]);
    }
  }
  double u[V][n], ub[V][n];
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int s = 0; s < n; ++s) {
      u[v][s] = (double)src[tbase + ((long long)v * g.NE + e) * NL + lb + s * sa];
      ub[v][s] = 0.0;
    }
  const Dual gam(prm.gamma);
#pragma unroll
  for (int fp = 1; fp < nf - 1; ++fp) {
    double M[V], Mb[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * u[v][s];
      M[v] = acc;
    }
    ad::vjp<V, V>(M, Fb[fp], Mb, [&](const Dual* xd, Dual* yd) {
      Dual W[V];
      prims<NDIM, Dual>(xd, W, a, gam);
      flux<NDIM, Dual>(W, yd, a, gam);
    });
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
      for (int s = 0; s < n; ++s) ub[v][s] += op.s2f[fp][s] * Mb[v];
  }
#pragma unroll
  for (int v = 0; v < V; ++v)
#pragma unroll
    for (int s = 0; s < n; ++s)
      ad::acc_add(&Sb[tbase + ((long long)v * g.NE + e) * NL + lb + s * sa], ub[v][s]);
}

// ---------------------------------------------------------------------------
// Picard update U_ad^(k)[tn] = U_sp - dt sum_p invader[tn][p] dudt[p]
// (src/sdader_simulator.py:187-200) transposed: Usb += sum_tn Uadb[tn],
// Db[p] = -dt sum_tn invader[tn][p] Uadb[tn]; with Tsrc == 1 (first iteration,
// one source node) Db[0] = sum_p Db[p].  mode 1: SD corrector
// U_sp' = U_sp - dt sum_t w_t dudt[t]: Db[t] = -dt w_t Usb (Usb unchanged).
// ---------------------------------------------------------------------------
template <int NDIM, int P>
__global__ void __launch_bounds__(256)
k_adj_picard(Ops<P> op, Geom g, int mode, int Tsrc, const double* __restrict__ dt_dev,
             const double* __restrict__ Uadb, double* __restrict__ Usb, double* __restrict__ Db) {
  constexpr int V = Dims<NDIM>::V, n = P + 1, T = P + 1;
  constexpr int NL = IPow<n, NDIM>::v;
  const long long npt = (long long)V * g.NE * NL;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= npt) return;
  const double dt = *dt_dev;
  if (mode == 1) {
    const double ub = Usb[tid];
#pragma unroll
    for (int t = 0; t < T; ++t) Db[t * npt + tid] = -dt * (op.w[t] * ub);
    return;
  }
  double ub[T];
  double s = Usb[tid];
#pragma unroll
  for (int tn = 0; tn < T; ++tn) {
    ub[tn] = Uadb[tn * npt + tid];
    s += ub[tn];
  }
  Usb[tid] = s;
  double tot = 0.0;
#pragma unroll
  for (int p = 0; p < T; ++p) {
    double acc = 0.0;
#pragma unroll
    for (int tn = 0; tn < T; ++tn) acc += op.inv[tn][p] * ub[tn];
    if (Tsrc == 1)
      tot += acc;
    else
      Db[p * npt + tid] = -dt * acc;
  }
  if (Tsrc == 1) Db[tid] = -dt * tot;
}

}  // namespace spd

// ===========================================================================
// Host side
// ===========================================================================

// Adjoint scratch of one context (FP64).
struct AdjScratch {
  double* Uadb = nullptr;       // (T, V, E, L) cotangent of a predictor iterate
  double* Db = nullptr;         // (T, V, E, L) cotangent of dudt
  double* Ffb[3] = {nullptr, nullptr, nullptr};  // (T, V, plane_a) element-face cotangents
  double* Qr[3] = {nullptr, nullptr, nullptr};   // (T, V, plane_a) right-trace cotangents
  double* Ucvb[2] = {nullptr, nullptr};          // (V, Nc) FV-state cotangents
  double* Fhob[3] = {nullptr, nullptr, nullptr};  // (V, nface_a) HO-flux cotangents
  double* WbL[3] = {nullptr, nullptr, nullptr};   // (V, nface_a) fallback left-state cotangents
  double* WbR[3] = {nullptr, nullptr, nullptr};
  double* HS = nullptr;         // ((1+d) V, (Ns+2)^d) Hancock states
  bool ready = false;
};

void spd_ctx::adj_release() {
  delete adj;
  adj = nullptr;
}

namespace {

using namespace spd;

long long plane_of(const spd_ctx* c, int a) {
  return (long long)(c->g.N[a] + 1) * (c->g.NE / c->g.N[a]) * c->NLP;
}
long long nface_of(const spd_ctx* c, int a) {
  long long nf = 1;
  for (int b = 0; b < c->ndim; ++b) nf *= c->g.Ns[b] + (b == a ? 1 : 0);
  return nf;
}
long long n1_of(const spd_ctx* c) {
  long long n1 = 1;
  for (int b = 0; b < c->ndim; ++b) n1 *= c->g.Ns[b] + 2;
  return n1;
}

int ensure_tape(spd_ctx* c) {
  if (c->tape_uad) return SPD_OK;
  const long long it = (long long)c->T * c->V * c->g.NE * c->NL;
  int rc;
  unsigned char* uad = nullptr;  // predictor iterates in the bulk type of the form
  const size_t bsz = c->cfg.precision == SPD_PRECISION_FP32 ? sizeof(float) : sizeof(double);
  if ((rc = c->alloc(&uad, (size_t)(c->cfg.m * it) * bsz)) != SPD_OK) return rc;
  c->tape_uad = uad;
  if (c->cfg.update == SPD_UPDATE_FV) {
    if ((rc = c->alloc(&c->tape_ucv, (size_t)(c->T * c->V * c->g.Nc))) != SPD_OK) return rc;
    if (c->cfg.FB && (rc = c->alloc(&c->tape_theta, (size_t)(c->T * c->g.Ngc))) != SPD_OK)
      return rc;
  }
  return SPD_OK;
}

// Adjoint scratch.  The reverse sweep reads the tape and never the forward
// step's own buffers, and the two never run at the same time (the taped
// forward of a step, then its reverse sweep), so in the FP64 form the
// cotangent arrays of the same size and type live in the forward's buffers:
// the iterate cotangent in the predictor state, the element-face cotangents
// in the element-face fluxes, the FV-face cotangents in the FV face fluxes,
// the Hancock states in the forward's (1-D / 3-D), and the two FV-state
// cotangents in the two FV state buffers (every form).  The rest is one pool
// shared by the two phases of the sweep: the FV substeps (fallback
// face-state cotangents, and what did not alias) and then the predictor
// (dudt cotangent, right-trace cotangents), which never hold values at the
// same time.  Within the predictor phase the dudt cotangent (read by the
// flux-point reverse kernels) and the right-trace cotangents (written by the
// element-face reverse kernels that follow) share their space too.
int ensure_scratch(spd_ctx* c) {
  if (!c->adj) c->adj = new AdjScratch();
  AdjScratch* A = c->adj;
  if (A->ready) return SPD_OK;
  const long long it = (long long)c->T * c->V * c->g.NE * c->NL;
  const bool f64 = c->cfg.precision == SPD_PRECISION_FP64;
  const bool fv = c->cfg.update == SPD_UPDATE_FV, fb = fv && c->cfg.FB;
  int rc;
  // phase 1 (FV substeps) and phase 2 (predictor) parts of the pool, in doubles
  long long p1 = 0, p2 = 0;
  const bool uadb_own = !(f64 && c->Uad);
  const bool ffb_own = !f64;
  const bool fhob_own = !(f64 && fv);
  const bool hs_own = !(f64 && c->HS);
  if (uadb_own) p1 += it;  // lives through both phases: first in the pool
  long long qr = 0;
  for (int a = 0; a < c->ndim; ++a) {
    if (ffb_own) p1 += (long long)c->T * c->V * plane_of(c, a);
    qr += (long long)c->T * c->V * plane_of(c, a);
  }
  const long long live = p1;  // Uadb and Ffb: both phases
  p2 = std::max(it, qr);      // Db, then Qr in the same space
  if (fv) {
    for (int a = 0; a < c->ndim; ++a) {
      const long long nfc = (long long)c->V * nface_of(c, a);
      if (fhob_own) p1 += nfc;
      if (fb) p1 += 2 * nfc;  // WbL, WbR
    }
    if (fb && hs_own) p1 += (long long)(1 + c->ndim) * c->V * n1_of(c);
    if (!c->Ucv) p1 += c->V * c->g.Nc;
    if (!c->Unew) p1 += c->V * c->g.Nc;
  }
  const long long pool_n = std::max(p1, live + p2);
  double* pool = nullptr;
  if ((rc = c->alloc(&pool, (size_t)pool_n)) != SPD_OK) return rc;
  double* q = pool;
  auto take = [&](long long n) {
    double* r = q;
    q += n;
    return r;
  };
  A->Uadb = uadb_own ? take(it) : (double*)c->Uad;
  for (int a = 0; a < c->ndim; ++a)
    A->Ffb[a] = ffb_own ? take((long long)c->T * c->V * plane_of(c, a)) : (double*)c->Fface[a];
  double* phase = q;  // both phases start here
  if (fv) {
    A->Ucvb[0] = c->Ucv ? c->Ucv : take(c->V * c->g.Nc);
    A->Ucvb[1] = c->Unew ? c->Unew : take(c->V * c->g.Nc);
    for (int a = 0; a < c->ndim; ++a) {
      const long long nfc = (long long)c->V * nface_of(c, a);
      A->Fhob[a] = fhob_own ? take(nfc) : (double*)c->Fho[a];
      if (fb) {
        A->WbL[a] = take(nfc);
        A->WbR[a] = take(nfc);
      }
    }
    if (fb) A->HS = hs_own ? take((long long)(1 + c->ndim) * c->V * n1_of(c)) : (double*)c->HS;
  }
  q = phase;
  A->Db = q;
  for (int a = 0; a < c->ndim; ++a) A->Qr[a] = take((long long)c->T * c->V * plane_of(c, a));
  for (int a = c->ndim; a < 3; ++a) {
    A->Ffb[a] = A->Ffb[0];
    A->Qr[a] = A->Qr[0];
    A->Fhob[a] = A->Fhob[0];
    A->WbL[a] = A->WbL[0];
    A->WbR[a] = A->WbR[0];
  }
  A->ready = true;
  return SPD_OK;
}

// B: element type of the tape's predictor iterates (the bulk type of the
// form: FP64, or FP32 in the FP32-bulk form); the reverse sweep itself is FP64.
template <int NDIM, int P, typename B> struct Adjoint {
  // blocks of k_adj_tensor3d (256 / (p+1)^3 elements each)
  static unsigned adj_t3d_blocks(long long NE) {
    constexpr int NL3 = (P + 1) * (P + 1) * (P + 1), EB = 256 / NL3 > 0 ? 256 / NL3 : 1;
    return (unsigned)((NE + EB - 1) / EB);
  }
  static constexpr int V = NDIM + 2, n = P + 1, nf = P + 2, T = P + 1;
  static constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;

  template <int A, typename S>
  static void face(spd_ctx* c, const Ops<P>& op, const S* src, int Tsrc, double* Sb,
                   hipStream_t s) {
    AdjScratch* Ad = c->adj;
    const long long tot = (long long)Tsrc * plane_of(c, A);
    if (c->cfg.riemann_sd != SPD_RS_LLF)
      hipLaunchKernelGGL((k_adj_face<NDIM, P, RS_HLLC, A, S>), dim3(grid_for(tot)), dim3(256), 0,
                         s, op, c->g, c->prm, src, Tsrc, Ad->Ffb[A], Ad->Qr[A]);
    else
      hipLaunchKernelGGL((k_adj_face<NDIM, P, RS_LLF, A, S>), dim3(grid_for(tot)), dim3(256), 0,
                         s, op, c->g, c->prm, src, Tsrc, Ad->Ffb[A], Ad->Qr[A]);
    const long long lines = (long long)Tsrc * c->g.NE * NLP;
    hipLaunchKernelGGL((k_adj_trace<NDIM, P, A>), dim3(grid_for(lines)), dim3(256), 0, s, op,
                       c->g, Tsrc, (const double*)Ad->Ffb[A], (const double*)Ad->Qr[A], Sb);
  }
  template <typename S>
  static void faces(spd_ctx* c, const Ops<P>& op, const S* src, int Tsrc, double* Sb,
                    hipStream_t s) {
    face<0>(c, op, src, Tsrc, Sb, s);
    if constexpr (NDIM >= 2) face<1>(c, op, src, Tsrc, Sb, s);
    if constexpr (NDIM >= 3) face<2>(c, op, src, Tsrc, Sb, s);
  }
  template <int A, typename S>
  static void pred_one(spd_ctx* c, const Ops<P>& op, const S* src, int Tsrc, double* Sb,
                       hipStream_t s) {
    const long long tot = (long long)Tsrc * c->g.NE * NLP;
    hipLaunchKernelGGL((k_adj_pred<NDIM, P, A, S>), dim3(grid_for(tot)), dim3(256), 0, s, op, c->g,
                       c->prm, src, Tsrc, c->adj->Db, Sb, c->adj->Ffb[A]);
  }
  // cotangent of D(src) (held in Db) -> Sb
  template <typename S>
  static void spatial(spd_ctx* c, const Ops<P>& op, const S* src, int Tsrc, double* Sb,
                      hipStream_t s) {
    for (int a = 0; a < NDIM; ++a)
      (void)hipMemsetAsync(c->adj->Ffb[a], 0, (size_t)Tsrc * V * plane_of(c, a) * sizeof(double),
                           s);
    pred_one<0>(c, op, src, Tsrc, Sb, s);
    if constexpr (NDIM >= 2) pred_one<1>(c, op, src, Tsrc, Sb, s);
    if constexpr (NDIM >= 3) pred_one<2>(c, op, src, Tsrc, Sb, s);
    faces(c, op, src, Tsrc, Sb, s);
  }
  template <int RS, int A>
  static void fvface(spd_ctx* c, const double* Ub1, const double* th, const double* dt, double w,
                     hipStream_t s) {
    AdjScratch* Ad = c->adj;
    hipLaunchKernelGGL((k_adj_fvface<NDIM, RS, A>), dim3(grid_for(nface_of(c, A))), dim3(256), 0,
                       s, c->g, c->prm, c->cfg.FB, Ub1, th, Ad->HS, dt, w, Ad->Fhob[A],
                       Ad->WbL[A], Ad->WbR[A]);
  }
  template <int RS>
  static void fvfaces(spd_ctx* c, const double* Ub1, const double* th, const double* dt, double w,
                      hipStream_t s) {
    fvface<RS, 0>(c, Ub1, th, dt, w, s);
    if constexpr (NDIM >= 2) fvface<RS, 1>(c, Ub1, th, dt, w, s);
    if constexpr (NDIM >= 3) fvface<RS, 2>(c, Ub1, th, dt, w, s);
  }
  // elements per block of k_adj_hoflux: 256 lines per dimension, a slab of at
  // most 40 KiB
  static int hof_epb() {
    return std::max(1, std::min(256 / NLP, 40960 / (V * NL * (int)sizeof(double))));
  }
  static void hoflux(spd_ctx* c, const Ops<P>& op, int i, const B* Uad, hipStream_t s) {
    AdjScratch* Ad = c->adj;
    const int E = hof_epb();
    hipLaunchKernelGGL((k_adj_hoflux<NDIM, P, B>), dim3((unsigned)((c->g.NE + E - 1) / E)),
                       dim3(256), (size_t)E * V * NL * sizeof(double), s, op, c->g, c->prm, i, Uad,
                       Ad->Fhob[0], Ad->Fhob[1], Ad->Fhob[2], Ad->Uadb, Ad->Ffb[0], Ad->Ffb[1],
                       Ad->Ffb[2], E);
  }

  static int backward(spd_ctx* c, const double* Usp, const double* dt, double* Ub,
                      hipStream_t s) {
    const Ops<P> op = make_ops<P>(*c->ops);
    AdjScratch* Ad = c->adj;
    const int m = c->cfg.m;
    const long long it_size = (long long)T * V * c->g.NE * NL;
    const long long npt = (long long)V * c->g.NE * NL;
    const B* tape = (const B*)c->tape_uad;
    auto iterate = [&](int k) { return tape + (long long)(k - 1) * it_size; };  // U_ad^(k)
    (void)hipMemsetAsync(Ad->Uadb, 0, (size_t)it_size * sizeof(double), s);
    if (c->cfg.update == SPD_UPDATE_FV) {
      const long long Nc = c->g.Nc;
      // U_sp' = cv_to_sp(U^T)
      if constexpr (NDIM == 3)
        hipLaunchKernelGGL((k_adj_tensor3d<P, 0>), dim3(adj_t3d_blocks(c->g.NE)), dim3(256), 0, s,
                           op, c->g, 0, (const double*)Ub, Ad->Ucvb[0]);
      else
        hipLaunchKernelGGL((k_adj_tensor<NDIM, P>), dim3(grid_for(npt)), dim3(256), 0, s, op,
                           c->g, 0, (const double*)Ub, 0, Ad->Ucvb[0], 1);
      int cur = 0;
      for (int i = T - 1; i >= 0; --i) {
        const double w = c->ops->w_tp[i];
        double* Ub1 = Ad->Ucvb[cur];
        double* Ub0 = Ad->Ucvb[1 - cur];
        const double* Ui = c->tape_ucv + (long long)i * V * Nc;
        const double* th = c->cfg.FB ? c->tape_theta + (long long)i * c->g.Ngc : nullptr;
        (void)hipMemcpyAsync(Ub0, Ub1, (size_t)V * Nc * sizeof(double), hipMemcpyDeviceToDevice,
                             s);
        if (c->cfg.FB)
          hipLaunchKernelGGL((k_adj_hs<NDIM>), dim3(grid_for(n1_of(c))), dim3(256), 0, s, c->g,
                             c->prm, Ui, th, dt, w, Ad->HS);
        if (c->cfg.riemann_fv != SPD_RS_LLF)
          fvfaces<RS_HLLC>(c, Ub1, th, dt, w, s);
        else
          fvfaces<RS_LLF>(c, Ub1, th, dt, w, s);
        if (c->cfg.FB)
          hipLaunchKernelGGL((k_adj_hancock<NDIM>), dim3(grid_for(n1_of(c))), dim3(256), 0, s,
                             c->g, c->prm, Ui, th, dt, w, Ad->WbL[0], Ad->WbL[1], Ad->WbL[2],
                             Ad->WbR[0], Ad->WbR[1], Ad->WbR[2], Ub0);
        hoflux(c, op, i, iterate(m), s);
        cur = 1 - cur;
      }
      // U^0 = sp_to_cv(U_sp)
      if constexpr (NDIM == 3)
        hipLaunchKernelGGL((k_adj_tensor3d<P, 1>), dim3(adj_t3d_blocks(c->g.NE)), dim3(256), 0, s,
                           op, c->g, 1, (const double*)Ad->Ucvb[cur], Ub);
      else
        hipLaunchKernelGGL((k_adj_tensor<NDIM, P>), dim3(grid_for(npt)), dim3(256), 0, s, op,
                           c->g, 1, (const double*)Ad->Ucvb[cur], 1, Ub, 0);
      // element faces of the final iterate (every node, written by k_adj_hoflux)
      faces(c, op, iterate(m), T, Ad->Uadb, s);
    } else {
      hipLaunchKernelGGL((k_adj_picard<NDIM, P>), dim3(grid_for(npt)), dim3(256), 0, s, op, c->g,
                         1, T, dt, (const double*)Ad->Uadb, Ub, Ad->Db);
      spatial(c, op, iterate(m), T, Ad->Uadb, s);
    }
    // Picard chain
    for (int k = m; k >= 1; --k) {
      const int Tsrc = (k == 1) ? 1 : T;
      hipLaunchKernelGGL((k_adj_picard<NDIM, P>), dim3(grid_for(npt)), dim3(256), 0, s, op, c->g,
                         0, Tsrc, dt, (const double*)Ad->Uadb, Ub, Ad->Db);
      if (k == 1) {
        spatial(c, op, Usp, 1, Ub, s);
      } else {
        (void)hipMemsetAsync(Ad->Uadb, 0, (size_t)it_size * sizeof(double), s);
        spatial(c, op, iterate(k - 1), T, Ad->Uadb, s);
      }
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, std::string("adjoint launch: ") +
                                                         hipGetErrorString(e));
  }
};

using BackFn = int (*)(spd_ctx*, const double*, const double*, double*, hipStream_t);
struct AdjEntry {
  int ndim, p, precision;
  BackFn back;
};
#define SPD_ADJ(D, P)                                                        \
  {D, P, SPD_PRECISION_FP64, &Adjoint<D, P, double>::backward},              \
      {D, P, SPD_PRECISION_FP32, &Adjoint<D, P, float>::backward}
const AdjEntry kAdj[] = {
    SPD_ADJ(1, 1), SPD_ADJ(1, 2), SPD_ADJ(1, 3), SPD_ADJ(1, 4), SPD_ADJ(1, 5),
    SPD_ADJ(2, 1), SPD_ADJ(2, 2), SPD_ADJ(2, 3), SPD_ADJ(2, 4), SPD_ADJ(2, 5),
    SPD_ADJ(3, 1), SPD_ADJ(3, 2), SPD_ADJ(3, 3),
};
#undef SPD_ADJ

int adj_check(const spd_ctx* c) {
  if (!c) return fail(SPD_E_INVALID, "null context");
  if (c->g.stored)
    return fail(SPD_E_UNSUPPORTED, "the adjoint is not built for \"ic\" / \"pressure\" sides");
  return SPD_OK;
}

}  // namespace

extern "C" {

int spd_adj_step(spd_ctx* ctx, const double* U_in, double* U_out, const double* dt_dev,
                 const double* theta_in, double* theta_out, void* stream) {
  int rc = adj_check(ctx);
  if (rc != SPD_OK) return rc;
  if (!U_in || !U_out || !dt_dev) return fail(SPD_E_INVALID, "null argument");
  if (U_in == U_out) return fail(SPD_E_INVALID, "U_in and U_out must differ");
  const bool fb = ctx->cfg.update == SPD_UPDATE_FV && ctx->cfg.FB;
  if ((theta_in || theta_out) && !fb)
    return fail(SPD_E_INVALID, "blending coefficients exist only with the FV fallback");
  if ((rc = ensure_tape(ctx)) != SPD_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  const size_t th_bytes = (size_t)ctx->T * ctx->g.Ngc * sizeof(double);
  SPD_HIP_CHECK(hipMemcpyAsync(U_out, U_in, (size_t)ctx->V * ctx->g.NE * ctx->NL * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
  if (theta_in)
    SPD_HIP_CHECK(hipMemcpyAsync(ctx->tape_theta, theta_in, th_bytes, hipMemcpyDeviceToDevice, s));
  ctx->tape_on = 1;
  ctx->tape_replay = theta_in ? 1 : 0;
  rc = spd_step(ctx, U_out, nullptr, nullptr, dt_dev, 0, stream);
  ctx->tape_on = 0;
  ctx->tape_replay = 0;
  if (rc != SPD_OK) return rc;
  if (theta_out)
    SPD_HIP_CHECK(hipMemcpyAsync(theta_out, ctx->tape_theta, th_bytes, hipMemcpyDeviceToDevice, s));
  return SPD_OK;
}

int spd_adj_backward(spd_ctx* ctx, const double* U_in, const double* dt_dev, double* U_bar,
                     void* stream) {
  int rc = adj_check(ctx);
  if (rc != SPD_OK) return rc;
  if (!U_in || !dt_dev || !U_bar) return fail(SPD_E_INVALID, "null argument");
  if (!ctx->tape_uad) return fail(SPD_E_INVALID, "no tape: call spd_adj_step first");
  if ((rc = ensure_scratch(ctx)) != SPD_OK) return rc;
  for (const AdjEntry& e : kAdj)
    if (e.ndim == ctx->ndim && e.p == ctx->p && e.precision == ctx->cfg.precision)
      return e.back(ctx, U_in, dt_dev, U_bar, (hipStream_t)stream);
  return fail(SPD_E_UNSUPPORTED, "no adjoint kernels for this dimension and order");
}

size_t spd_adj_theta_count(const spd_ctx* ctx) {
  if (!ctx || ctx->cfg.update != SPD_UPDATE_FV || !ctx->cfg.FB) return 0;
  return (size_t)ctx->T * ctx->g.Ngc;
}

}  // extern "C"
