// Device-side building blocks of the SD-ADER step: pointwise Euler algebra,
// the two Riemann solvers, NaN-propagating extrema and the index arithmetic
// of the element and finite-volume layouts.  The pointwise functions take the
// arithmetic type of their caller: FP64, or FP32 in the FP32-bulk form.
//
// Variable order (both sets): 0 = density, 1..NDIM = momentum / velocity
// components (x first), NDIM+1 = total energy / pressure.  A "rolled" velocity
// list for dimension `a` starts with the normal component a+1 and continues
// cyclically, which is `np.roll(vels, -a)` in the reference
// (src/sdader_simulator.py:219, src/fv_simulator.py:149).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace spd {

enum : int { RS_LLF = 0, RS_HLLC = 1 };
enum : int { BC_PERIODIC = 0, BC_GRADFREE = 1, BC_REFLECTIVE = 2, BC_IC = 3, BC_PRESSURE = 4 };
// Sides whose boundary values are stored at construction ("ic", "pressure").
__host__ __device__ __forceinline__ bool bc_stored(int kind) { return kind >= BC_IC; }

constexpr int NGH = 2;  // finite-volume ghost layers (Nghc in the reference)

template <int NDIM> struct Dims {
  static constexpr int V = NDIM + 2;
  static constexpr int P = NDIM + 1;  // index of energy / pressure
};

template <int B, int E> struct IPow { static constexpr int v = B * IPow<B, E - 1>::v; };
template <int B> struct IPow<B, 0> { static constexpr int v = 1; };

// Bulk arithmetic type of the FP32 form.  The FP32 form uses the hardware
// reciprocal and square root (1 ulp).  The FP64 form divides by multiplying
// with the hardware reciprocal refined by two Newton steps (the quotient
// within an ulp; the correctly rounded division sequence costs more
// instructions: the divisions were a sixth of the predictor's) and keeps the
// correctly rounded square root.  Other types (the adjoint's dual numbers)
// divide.
template <typename R> struct IsF32 { static constexpr bool v = false; };
template <> struct IsF32<float> { static constexpr bool v = true; };

// 1 / b in FP64: v_rcp_f64 and two Newton steps (correctly rounded on every
// one of 6.7e7 log-uniform samples over [1e-6, 1e6]); where the hardware
// value is 0, infinite or NaN (b infinite, zero, subnormal or NaN) it is the
// result.
__device__ __forceinline__ double nrcp(double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  if (!(__builtin_isfinite(r0) && r0 != 0.0)) return r0;
  const double e0 = __builtin_fma(-b, r0, 1.0);
  const double r1 = __builtin_fma(r0, e0, r0);
  const double e1 = __builtin_fma(-b, r1, 1.0);
  return __builtin_fma(r1, e1, r1);
}
template <typename R> __device__ __forceinline__ R qdiv(R a, R b) {
  if constexpr (IsF32<R>::v)
    return a * __builtin_amdgcn_rcpf(b);
  else if constexpr (std::is_same<R, double>::value)
    return a * nrcp(b);
  else
    return a / b;
}
template <typename R> __device__ __forceinline__ R qrcp(R b) {
  if constexpr (IsF32<R>::v)
    return __builtin_amdgcn_rcpf(b);
  else if constexpr (std::is_same<R, double>::value)
    return nrcp(b);
  else
    return R(1) / b;
}
template <typename R> __device__ __forceinline__ R qsqrt(R x) {
  if constexpr (IsF32<R>::v)
    return __builtin_amdgcn_sqrtf(x);
  else
    return sqrt(x);
}

// 1 / (gamma - 1): the same value at every call of a thread, so the compiler
// evaluates it once; multiplying by it is within an ulp of the division of
// src/hydro.py:118-124,159-166.
template <typename R> __device__ __forceinline__ R inv_gm1(R gamma) {
  return qrcp<R>(gamma - R(1));
}

// numpy.maximum / numpy.minimum: NaN in either operand propagates.
template <typename R> __device__ __forceinline__ R nmax(R a, R b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}
template <typename R> __device__ __forceinline__ R nmin(R a, R b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? a : b;
}

// Rolled velocity index k (0-based position in the rolled list) for dim a.
template <int NDIM> __device__ __forceinline__ constexpr int rolled(int a, int k) {
  return 1 + (a + k) % NDIM;
}

// Evaluation without contraction into fused multiply-adds.  The conservative
// form, sound speed, Riemann solvers and moncen limiter below always round
// each operation on its own; the primitives and the Euler flux come in two
// forms, prims / flux (contracted: the predictor's volume terms, the faces,
// the detection) and prims_rn / flux_rn (each operation rounded).  Every
// kernel that inlines an uncontracted formula forms the same bits from the
// same inputs: the 2-D fallback's column-strip and compacted-tile paths take
// the high-order face fluxes, the old primitives and the fallback fluxes from
// these forms and give the same results.  Their rounding is that of the
// reference's NumPy expressions.

// conservative -> primitive (src/hydro.py:48-86), rolled list of dim a.  The
// FP32 form recovers the pressure from the total energy in FP64.  The FP64
// form takes one reciprocal of the density for all velocity components
// (within an ulp of the per-component divisions of src/hydro.py:79-81).
#define SPD_PRIMS_BODY                                                   \
  constexpr int P = Dims<NDIM>::P;                                       \
  W[0] = U[0];                                                           \
  if constexpr (IsF32<R>::v) {                                           \
    const float r = __builtin_amdgcn_rcpf(U[0]);                         \
    double K = 0.0;                                                      \
    _Pragma("unroll") for (int k = 0; k < NDIM; ++k) {                   \
      const int v = rolled<NDIM>(a, k);                                  \
      W[v] = U[v] * r;                                                   \
      K = K + (double)U[v] * (double)W[v];                               \
    }                                                                    \
    W[P] = (float)(((double)gamma - 1.0) * ((double)U[P] - 0.5 * K));    \
  } else {                                                               \
    const R rr = qrcp<R>(U[0]);                                          \
    R K = U[0] * R(0.0);                                                 \
    _Pragma("unroll") for (int k = 0; k < NDIM; ++k) {                   \
      const int v = rolled<NDIM>(a, k);                                  \
      W[v] = U[v] * rr;                                                  \
      K = K + W[v] * W[v];                                               \
    }                                                                    \
    K = K * (R(0.5) * U[0]);                                             \
    W[P] = (gamma - R(1.0)) * (U[P] - K);                                \
  }
template <int NDIM, typename R>
__device__ __forceinline__ void prims(const R* U, R* W, int a, R gamma) {
  SPD_PRIMS_BODY
}
template <int NDIM, typename R>
__device__ __forceinline__ void prims_rn(const R* U, R* W, int a, R gamma) {
#pragma clang fp contract(off)
  SPD_PRIMS_BODY
}
#undef SPD_PRIMS_BODY

// primitive -> conservative (src/hydro.py:88-125), rolled list of dim a.
template <int NDIM, typename R>
__device__ __forceinline__ void cons(const R* W, R* U, int a, R gamma) {
#pragma clang fp contract(off)
  constexpr int P = Dims<NDIM>::P;
  U[0] = W[0];
  R K = W[0] * R(0);
#pragma unroll
  for (int k = 0; k < NDIM; ++k) {
    const int v = rolled<NDIM>(a, k);
    U[v] = W[v] * U[0];
    K = K + W[v] * W[v];
  }
  K = K * (R(0.5) * U[0]);
  U[P] = W[P] * inv_gm1<R>(gamma) + K;
}

// Euler flux along dim a from primitives (src/hydro.py:127-170): the kinetic
// sum runs over the rolled list in reverse, so the normal component is last.
#define SPD_FLUX_BODY                                                    \
  constexpr int P = Dims<NDIM>::P;                                       \
  const int v1 = a + 1;                                                  \
  R K = W[0] * R(0);                                                     \
  R m = R(0);                                                            \
  _Pragma("unroll") for (int k = NDIM - 1; k >= 0; --k) {                \
    const int v = rolled<NDIM>(a, k);                                    \
    m = W[0] * W[v];                                                     \
    K = K + m * W[v];                                                    \
    F[v] = m * W[v1];                                                    \
  }                                                                      \
  const R E = W[P] * inv_gm1<R>(gamma) + R(0.5) * K;                     \
  F[0] = m;                                                              \
  F[v1] = m * W[v1] + W[P];                                              \
  F[P] = W[v1] * (E + W[P]);
template <int NDIM, typename R>
__device__ __forceinline__ void flux(const R* W, R* F, int a, R gamma) {
  SPD_FLUX_BODY
}
template <int NDIM, typename R>
__device__ __forceinline__ void flux_rn(const R* W, R* F, int a, R gamma) {
#pragma clang fp contract(off)
  SPD_FLUX_BODY
}
#undef SPD_FLUX_BODY

// Sound speed with the where-clamp of src/hydro.py:3-46 (NaN -> min_c2).
template <typename R> __device__ __forceinline__ R sound(R P, R rho, R gamma, R min_c2) {
#pragma clang fp contract(off)
  R c2 = qdiv<R>(gamma * P, rho);
  c2 = (c2 > min_c2) ? c2 : min_c2;
  return qsqrt<R>(c2);
}

// Local Lax-Friedrichs flux (src/riemann_solver.py:31-65).
template <int NDIM, typename R>
__device__ __forceinline__ void llf(const R* WL, const R* WR, const R* UL, const R* UR, R* F,
                                    int a, R gamma, R min_c2) {
#pragma clang fp contract(off)
  constexpr int V = Dims<NDIM>::V;
  constexpr int P = Dims<NDIM>::P;
  const int v1 = a + 1;
  R FL[V], FR[V];
  flux_rn<NDIM, R>(WL, FL, a, gamma);
  flux_rn<NDIM, R>(WR, FR, a, gamma);
  const R cL = sound<R>(WL[P], WL[0], gamma, min_c2) + fabs(WL[v1]);
  const R cR = sound<R>(WR[P], WR[0], gamma, min_c2) + fabs(WR[v1]);
  const R cmax = (cL > cR) ? cL : cR;
#pragma unroll
  for (int v = 0; v < V; ++v)
    F[v] = R(0.5) * (FR[v] + FL[v]) - R(0.5) * cmax * (UR[v] - UL[v]);
}

// HLLC flux in the Godunov-state form of src/riemann_solver.py:67-145.  Only
// the selected star state is evaluated; its formula is the reference's.  With
// low_mach the lhllc variant (src/riemann_solver.py:147-212): the velocity
// jump term of the star pressure is scaled by phi = M (2 - M),
// M = min(1, max(|v_L|, |v_R|) / c_max).
template <int NDIM, typename R>
__device__ __forceinline__ void hllc(const R* WL, const R* WR, const R* UL, const R* UR, R* F,
                                     int a, R gamma, R min_c2, int low_mach) {
#pragma clang fp contract(off)
  constexpr int P = Dims<NDIM>::P;
  const int v1 = a + 1;
  const R cL = sound<R>(WL[P], WL[0], gamma, min_c2) + fabs(WL[v1]);
  const R cR = sound<R>(WR[P], WR[0], gamma, min_c2) + fabs(WR[v1]);
  const R cmax = (cL > cR) ? cL : cR;
  const R vL = WL[v1], vR = WR[v1];
  const R sL = ((vL < vR) ? vL : vR) - cmax;
  const R sR = ((vL > vR) ? vL : vR) + cmax;
  const R rcL = WL[0] * (vL - sL);
  const R rcR = WR[0] * (sR - vR);
  // one reciprocal per shared denominator (differs from two divisions by at
  // most an ulp per value)
  const R rden = qrcp<R>(rcR + rcL);
  const R vs = (rcR * vR + rcL * vL + (WL[P] - WR[P])) * rden;
  R jump = rcL * rcR * (vL - vR);
  if (low_mach) {
    const R aL = fabs(vL), aR = fabs(vR);
    const R mach = nmin(R(1), qdiv<R>(nmax(aL, aR), cmax));
    jump = mach * (R(2) - mach) * rcL * rcR * (vL - vR);
  }
  const R Ps = (rcR * WL[P] + rcL * WR[P] + jump) * rden;
  R r, v, p, e;
  if (sL > R(0)) {
    r = WL[0]; v = vL; p = WL[P]; e = UL[P];
  } else if (vs > R(0)) {
    const R rs = qrcp<R>(sL - vs);
    r = WL[0] * (sL - vL) * rs;
    e = ((sL - vL) * UL[P] - WL[P] * vL + Ps * vs) * rs;
    v = vs; p = Ps;
  } else if (sR > R(0)) {
    const R rs = qrcp<R>(sR - vs);
    r = WR[0] * (sR - vR) * rs;
    e = ((sR - vR) * UR[P] - WR[P] * vR + Ps * vs) * rs;
    v = vs; p = Ps;
  } else {
    r = WR[0]; v = vR; p = WR[P]; e = UR[P];
  }
  F[0] = r * v;
  F[v1] = F[0] * v + p;
  F[P] = v * (e + p);
#pragma unroll
  for (int k = 1; k < NDIM; ++k) {
    const int vt = rolled<NDIM>(a, k);
    F[vt] = F[0] * ((vs > R(0)) ? WL[vt] : WR[vt]);
  }
}

// RS_HLLC covers hllc and lhllc (low_mach != 0).
template <int NDIM, int RS, typename R>
__device__ __forceinline__ void riemann(const R* WL, const R* WR, const R* UL, const R* UR, R* F,
                                        int a, R gamma, R min_c2, int low_mach) {
  if constexpr (RS == RS_HLLC)
    hllc<NDIM, R>(WL, WR, UL, UR, F, a, gamma, min_c2, low_mach);
  else
    llf<NDIM, R>(WL, WR, UL, UR, F, a, gamma, min_c2);
}

// Slope limiters of src/muscl.py:12-59 applied to the one-sided slopes dL, dR
// of a cell; xL, xR = centre spacings, xM = cell width (moncen only).
// minmod is (R/L clipped to [., 1], then to [0, .]) * L (Q7).  Evaluated
// without the division: same-sign slopes give R when |R| < |L| (the
// reference's (R/L) * L, within an ulp) and L otherwise; opposite signs or a
// zero give 0; a NaN R gives L and a NaN L gives NaN, as the ratio rules do.
template <typename R> __device__ __forceinline__ R minmod(R dL, R dR) {
  if (dL != dL || dR != dR) return dL;
  if (!(dL * dR > R(0))) return dL * R(0);
  return (fabs(dR) < fabs(dL)) ? dR : dL;
}
template <typename R> __device__ __forceinline__ R moncen(R dL, R dR, R xL, R xR, R xM) {
#pragma clang fp contract(off)
  const R dC = qdiv<R>(xL * dL + xR * dR, xL + xR);
  R slope = nmin(fabs(qdiv<R>(R(2) * dL * xL, xM)), fabs(qdiv<R>(R(2) * dR * xR, xM)));
  const R sg = (dC > R(0)) ? R(1) : ((dC < R(0)) ? R(-1) : dC);  // numpy.sign
  slope = sg * nmin(slope, R(fabs(dC)));
  return (dL * dR >= R(0)) ? slope : R(0);
}

// ---------------------------------------------------------------------------
// Problem description shared by all kernels (passed by value).
// ---------------------------------------------------------------------------
struct Geom {
  int N[3];       // elements per dimension (x, y, z); 1 for absent dims
  int Ns[3];      // subcells per dimension = N * (p+1)
  int Ng[3];      // ghosted subcells per dimension = Ns + 2*NGH
  int bc[3][2];   // boundary kind per dimension and side
  double h[3];    // element widths
  // subcell geometry of the ghosted grid, per dimension
  const double* faces[3];    // Ns + 5 face positions
  const double* centers[3];  // Ns + 4 cell centres
  const double* hfp[3];      // Ns + 4 cell widths
  const double* hcv[3];      // Ns + 3 centre spacings
  const double* rhfp[3];     // 1 / hfp
  const double* rhcv[3];     // 1 / hcv
  const double* rdc[3];      // 1 / (centers[k+1] - centers[k-1]), k in [1, Ns+2]
  double rh[3];              // 1 / h
  const int* skip;  // device flag: when non-zero every step kernel returns at once
  // stored boundary values of "ic" / "pressure" sides (null otherwise):
  // element traces per dimension and side, (V, points across the side), and
  // the initial primitives on the ghosted FV grid, (V, Ngc)
  const double* bct[3][2];
  const double* wgh;
  int stored;     // some side keeps stored values
  long long NE;   // number of elements
  long long Nc;   // number of active subcells
  long long Ngc;  // number of ghosted subcells
};

#define SPD_SKIP(g) \
  if (*(g).skip) return

// Tile of this block of a 2-D tile grid, with the tiles dealt to the 8 XCDs in
// column bands.  The dispatcher hands consecutive workgroups round-robin to the
// XCDs, each with its own L2, so in launch order the tiles left and right of a
// tile run on other XCDs and every x halo is fetched again.  With gridDim.x a
// multiple of 8, XCD k takes the tile columns [k W, (k + 1) W), W = gridDim.x
// / 8, in row order: a tile's neighbours run on its XCD one and W blocks
// apart.  Other grids keep the launch order.
// Block index of a launch (x, y) in a grid of nx x ny blocks: the hardware's,
// or inside a batched launch (one launch for several simulations) the
// block's place in the grid of the simulation it works for.
struct BlockIx {
  int x, y, nx, ny;
};
__device__ __forceinline__ BlockIx hw_block() {
  return BlockIx{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y};
}

__device__ __forceinline__ void xcd_band_tile(const BlockIx& bi, int& bx, int& by) {
  const int nbx = bi.nx;
  if (nbx % 8 != 0) {
    bx = bi.x;
    by = bi.y;
    return;
  }
  const int b = bi.y * nbx + bi.x;
  const int W = nbx >> 3, j = b >> 3;
  bx = (b & 7) * W + j % W;
  by = j / W;
}
__device__ __forceinline__ void xcd_band_tile(int& bx, int& by) {
  xcd_band_tile(hw_block(), bx, by);
}

struct Params {
  double gamma, min_c2;
  double tol, min_rho, max_rho, min_P;
  int p;
  int nad_delta, sed, pad, blending, godunov;
  int lowmach_sd, lowmach_fv;  // lhllc on the element / fallback faces
  int moncen;                  // slope limiter: 0 minmod, 1 moncen
};

// Map a ghosted-grid coordinate (active range [0, Ns)) to the active cell whose
// value the finite-volume ghost fill copies there (src/fv_simulator.py:224-273):
// periodic wrap, or the non-mirrored copy of the first / last NGH active layers.
// Returns the mapped coordinate and sets *flip when a reflective side negates
// the normal velocity.
__device__ __forceinline__ int ghost_map(int c, int Ns, const int* bc, bool* flip) {
  if (c < 0) {
    if (bc[0] == BC_PERIODIC) return c + Ns;
    if (bc[0] == BC_REFLECTIVE) *flip = !*flip;
    return c + NGH;
  }
  if (c >= Ns) {
    if (bc[1] == BC_PERIODIC) return c - Ns;
    if (bc[1] == BC_REFLECTIVE) *flip = !*flip;
    return c - NGH;
  }
  return c;
}

// FV-layout linear index of an active cell (x fastest).
template <int NDIM>
__device__ __forceinline__ long long cell_index(const Geom& g, const int* c) {
  long long idx = 0;
#pragma unroll
  for (int a = NDIM - 1; a >= 0; --a) idx = idx * g.Ns[a] + c[a];
  return idx;
}

// Solution point across dimension a of the stored element traces: the
// transverse coordinates (element ce[b], line cl[b]) of the other
// dimensions, higher dimensions slowest (the face grids' order).
template <int NDIM>
__device__ __forceinline__ long long bc_point(const Geom& g, int a, int n, const int* ce,
                                              const int* cl) {
  long long t = 0;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b)
    if (b != a) t = t * g.Ns[b] + (long long)ce[b] * n + cl[b];
  return t;
}

// Element-face traces at a domain face of dimension a whose side (0: the
// first face, 1: the last) keeps stored values: the outer state becomes the
// trace stored at construction, and a "pressure" side also overwrites the
// boundary element's own trace with it (src/sd_boundary.py:50-56).
template <int V, typename B>
__device__ __forceinline__ void bc_stored_traces(const Geom& g, int a, int side, long long tp,
                                                 B* outer, B* inner) {
  const double* st = g.bct[a][side];
  const long long P = g.Nc / g.Ns[a];
#pragma unroll
  for (int v = 0; v < V; ++v) outer[v] = (B)st[v * P + tp];
  if (g.bc[a][side] == BC_PRESSURE) {
#pragma unroll
    for (int v = 0; v < V; ++v) inner[v] = outer[v];
  }
}

// Ghosted-grid index of coordinate c (each in [-NGH, Ns + NGH)).
template <int NDIM>
__device__ __forceinline__ long long ghosted_index(const Geom& g, const int* c) {
  long long idx = 0;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) idx = idx * g.Ng[b] + c[b] + NGH;
  return idx;
}

// FV ghost value at a ghosted coordinate c that the stored ghost layers of
// an "ic" / "pressure" side provide (the reference never refreshes them,
// src/fv_simulator.py:213-253): the ghost fill runs x first, so the
// dimensions are resolved from the last: a ghost coordinate of a side that
// copies active cells is mapped (periodic wrap, non-mirrored copy, reflective
// flip), and the first stored side met takes the value of its layers at the
// coordinates reached, the lower dimensions unmapped.  Returns false (W
// untouched) when no stored side is met.
template <int NDIM>
__device__ __forceinline__ bool stored_ghost(const Geom& g, const int* c, double* W) {
  if (!g.stored) return false;
  constexpr int V = NDIM + 2;
  int m[3] = {c[0], c[1], c[2]};
  bool flip[3] = {false, false, false};
#pragma unroll
  for (int a = NDIM - 1; a >= 0; --a) {
    if (c[a] < 0 || c[a] >= g.Ns[a]) {
      const int side = (c[a] < 0) ? 0 : 1;
      if (bc_stored(g.bc[a][side])) {
        const long long idx = ghosted_index<NDIM>(g, m);
#pragma unroll
        for (int v = 0; v < V; ++v) W[v] = g.wgh[v * g.Ngc + idx];
#pragma unroll
        for (int b = a + 1; b < NDIM; ++b)
          if (flip[b]) W[b + 1] = -W[b + 1];
        return true;
      }
      m[a] = ghost_map(c[a], g.Ns[a], g.bc[a], &flip[a]);
    }
  }
  return false;
}

// Primitive state of the ghost-filled array at a ghosted coordinate `c`
// (active range [0, Ns) per dimension; ghosts in [-NGH, 0) and [Ns, Ns+NGH)).
// The value is the one left by the fill x first, then y, then z: every ghost
// coordinate is mapped independently and each reflective crossing negates its
// own velocity component.
template <int NDIM>
__device__ __forceinline__ void load_W(const double* __restrict__ U, const Geom& g, const int* c,
                                       double* W, double gamma) {
  constexpr int V = Dims<NDIM>::V;
  if (stored_ghost<NDIM>(g, c, W)) return;
  int m[3];
  bool flip[3] = {false, false, false};
#pragma unroll
  for (int a = 0; a < NDIM; ++a) m[a] = ghost_map(c[a], g.Ns[a], g.bc[a], &flip[a]);
  const long long idx = cell_index<NDIM>(g, m);
  double u[V];
#pragma unroll
  for (int v = 0; v < V; ++v) u[v] = U[v * g.Nc + idx];
  prims_rn<NDIM>(u, W, 0, gamma);
#pragma unroll
  for (int a = 0; a < NDIM; ++a)
    if (flip[a]) W[a + 1] = -W[a + 1];
}

// The same value from a precomputed primitive array Wp (V, Nc) of the state.
template <int NDIM>
__device__ __forceinline__ void load_prim(const double* __restrict__ Wp, const Geom& g,
                                          const int* c, double* W) {
  constexpr int V = Dims<NDIM>::V;
  if (stored_ghost<NDIM>(g, c, W)) return;
  int m[3];
  bool flip[3] = {false, false, false};
#pragma unroll
  for (int a = 0; a < NDIM; ++a) m[a] = ghost_map(c[a], g.Ns[a], g.bc[a], &flip[a]);
  const long long idx = cell_index<NDIM>(g, m);
#pragma unroll
  for (int v = 0; v < V; ++v) W[v] = Wp[v * g.Nc + idx];
#pragma unroll
  for (int a = 0; a < NDIM; ++a)
    if (flip[a]) W[a + 1] = -W[a + 1];
}

// Limited slope of one variable along dimension b of the cell at coordinate
// c (src/muscl.py:61-75; Q7): from the cell's value w and its neighbours'
// wm (c - e_b) and wp (c + e_b).
template <typename R>
__device__ __forceinline__ R hancock_slope(const Geom& g, const Params& prm, int b, int cb, R w,
                                           R wm, R wp) {
  const int G = cb + NGH;
  const R L = (w - wm) * (R)g.rhcv[b][G - 1];
  const R Rr = (wp - w) * (R)g.rhcv[b][G];
  return prm.moncen ? moncen<R>(L, Rr, (R)g.hcv[b][G - 1], (R)g.hcv[b][G], (R)g.hfp[b][G])
                    : minmod<R>(L, Rr);
}

// Half-step predicted state from the primitives W and the limited slopes
// dMh[b] (src/muscl.py:129-173; Q1: pressure row zeroed, Q19: the transverse
// loop over the unrolled velocity list).
template <int NDIM, typename R>
__device__ __forceinline__ void hancock_predict(const Params& prm, const R* W,
                                                const R (*dMh)[Dims<NDIM>::V], R dti, R* Mp) {
  constexpr int V = Dims<NDIM>::V, P = Dims<NDIM>::P;
  const R gam = (R)prm.gamma;
  R dtW[V];
#pragma unroll
  for (int v = 0; v < V; ++v) dtW[v] = R(0);
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    const int vel = b + 1;
    const R* dW = dMh[b];
    dtW[0] = dtW[0] - (W[vel] * dW[0] + W[0] * dW[vel]);
    dtW[P] = dtW[P] - (W[vel] * dW[P] + gam * W[P] * dW[vel]);
    dtW[vel] = dtW[vel] - (W[vel] * dW[vel] + qdiv<R>(dW[P], W[0]));
#pragma unroll
    for (int v2 = 2; v2 <= NDIM; ++v2) dtW[v2] = dtW[v2] - W[vel] * dW[v2];
  }
  dtW[P] = dtW[P] * R(0);
#pragma unroll
  for (int v = 0; v < V; ++v) Mp[v] = W[v] + R(0.5) * dtW[v] * dti;
}

// MUSCL-Hancock state of one cell of the one-layer-ghosted grid
// (src/muscl.py:61-75,129-221): limited slopes of the primitives W from the
// neighbours Wm[b] (c - e_b) and Wp[b] (c + e_b), the half slopes
// S[b] = slope * width / 2 and the half-step predicted state Mp.  `c` is the
// cell's coordinate (active range [0, Ns)).
template <int NDIM, typename R>
__device__ __forceinline__ void hancock_state(const Geom& g, const Params& prm, const int* c,
                                              const R* W, const R (*Wm)[Dims<NDIM>::V],
                                              const R (*Wp)[Dims<NDIM>::V], R dti, R* Mp,
                                              R (*S)[Dims<NDIM>::V]) {
  constexpr int V = Dims<NDIM>::V;
  R dMh[NDIM][V];
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {
    const R hf = (R)g.hfp[b][c[b] + NGH];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      dMh[b][v] = hancock_slope<R>(g, prm, b, c[b], W[v], Wm[b][v], Wp[b][v]);
      S[b][v] = R(0.5) * dMh[b][v] * hf;
    }
  }
  hancock_predict<NDIM, R>(prm, W, dMh, dti, Mp);
}

}  // namespace spd
