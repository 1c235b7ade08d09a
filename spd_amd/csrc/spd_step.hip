// MI355X (gfx950) kernels of the SD-ADER step and the C ABI of libspd_hip.so.
//
// Data layouts (FP64; in the FP32-bulk form the predictor state, the element-
// face fluxes, the FV-face fluxes and the Hancock states are FP32):
//   element layout   (V, E, L)        E = elements (x fastest), L = (p+1)^d points
//   predictor state  (T, V, E, L)     T = m+1 time nodes
//   element faces    (T, V, N_a+1, E_perp, L_perp) per dimension a
//   FV cells         (V, Ns_z, Ns_y, Ns_x)
//   FV faces         (V, ...) with Ns_a+1 faces along dimension a
//   ghosted FV grid  (Ns_z+4, Ns_y+4, Ns_x+4)  (theta)
//
// One step (SDADER_Simulator.perform_update, src/sdader_simulator.py:361-378):
//   predictor  m+1 x [element-face Riemann (k_face_sd) + update (k_pred)]
//              (2-D: m x k_pred2d, element faces solved inside, then the
//              faces of the final state for the corrector in k_face_sd2d)
//   corrector  SD : k_pred in SD-update mode
//              FV : sp->cv (k_tensor2d / k_tensor3d / k_tensor_el), then T substeps of
//                   [HO face fluxes (k_hoflux), candidate (k_apply),
//                    detection (k_detect), theta (k_theta1, k_ring),
//                    fallback states (k_hancock) + flux and blend (k_fbface),
//                    apply (k_apply)],
//                   then cv->sp with the wave speed for the next dt
//                   (k_tensor2d / k_tensor3d; 1-D: k_tensor_el + k_wave_cells).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spd.h"
#include "spd_device.hpp"
#include "spd_host_ops.hpp"
#include "spd_layout.hpp"
#include "spd_ctx.hpp"

namespace spd {

// Dynamic LDS of the templated kernels (typed at each use).
extern __shared__ __attribute__((aligned(16))) unsigned char spd_dyn_lds[];

// Maximum over the block (NaN-propagating); the result is valid in thread 0.
__device__ __forceinline__ double block_max(double v) {
  __shared__ double red[32];
  for (int off = warpSize / 2; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off));
  const int w = threadIdx.x / warpSize, lane = threadIdx.x % warpSize;
  const int nw = (blockDim.x + warpSize - 1) / warpSize;
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (w == 0) {
    v = (lane < nw) ? red[lane] : 0.0;
    for (int off = warpSize / 2; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off));
  }
  return v;
}

// ---------------------------------------------------------------------------
// Predictor: Riemann problems at element faces for every time node
// (src/sdader_simulator.py:213-240, src/sd_boundary.py:7-76).
// One thread per (t, face, perpendicular element, perpendicular line).
// ---------------------------------------------------------------------------
// x-faces (a = 0) with NDIM >= 2: a block takes FPB consecutive faces of all
// NLP transverse lines of one element row, so that the lanes reading one
// element's rows are adjacent; the fluxes are transposed through LDS and
// written face-innermost.
template <int NDIM, int P, int RS, typename S, typename B>
__global__ void __launch_bounds__(256)
k_face_sd_x(Ops<P, B> op, Geom g, Params prm, const S* __restrict__ src, int Tsrc,
            B* __restrict__ Fface) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;
  constexpr int FPB = 256 / NLP;
  __shared__ B buf[V][NLP][FPB];
  const int Na = g.N[0];
  const int nchunk = (Na + 1 + FPB - 1) / FPB;
  const long long NEp = g.NE / Na;
  long long b = blockIdx.x;
  const int fc = (int)(b % nchunk);
  b /= nchunk;
  const long long ep = b % NEp;
  const int t = (int)(b / NEp);
  const int lp = threadIdx.x % NLP, fi = threadIdx.x / NLP;
  const int f = fc * FPB + fi;
  int ce[3] = {0, 0, 0}, cl[3] = {0, 0, 0};
  {
    long long x = ep;
#pragma unroll
    for (int d = 1; d < NDIM; ++d) {
      ce[d] = split_index(x, g.N[d]);
    }
    int r = lp;
#pragma unroll
    for (int d = 1; d < NDIM; ++d) {
      cl[d] = r % n;
      r /= n;
    }
  }
  int lb = 0;  // local index of point 0 of the line (x coordinate 0)
#pragma unroll
  for (int d = NDIM - 1; d >= 1; --d) lb = lb * n + cl[d];
  lb *= n;
  const S* base = src + (long long)t * V * g.NE * NL;
  auto trace = [&](int ea, int fp, B* q) {
    int c[3] = {ea, ce[1], ce[2]};
    const long long e = elem_index<NDIM>(g, c);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const S* u = base + ((long long)v * g.NE + e) * NL + lb;
      B acc = B(0);
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * (B)u[s];
      q[v] = acc;
    }
  };
  if (fi < FPB && f <= Na) {
    B qL[V], qR[V];
    if (f > 0) trace(f - 1, nf - 1, qL);
    if (f < Na) trace(f, 0, qR);
    if (f == 0) {
      const int k = g.bc[0][0];
      if (k == BC_PERIODIC) {
        trace(Na - 1, nf - 1, qL);
      } else if (bc_stored(k)) {
        bc_stored_traces<V, B>(g, 0, 0, bc_point<NDIM>(g, 0, n, ce, cl), qL, qR);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) qL[v] = qR[v];
        if (k == BC_REFLECTIVE) qL[1] = -qL[1];
      }
    }
    if (f == Na) {
      const int k = g.bc[0][1];
      if (k == BC_PERIODIC) {
        trace(0, 0, qR);
      } else if (bc_stored(k)) {
        bc_stored_traces<V, B>(g, 0, 1, bc_point<NDIM>(g, 0, n, ce, cl), qR, qL);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) qR[v] = qL[v];
        if (k == BC_REFLECTIVE) qR[1] = -qR[1];
      }
    }
    B WL[V], WR[V], F[V];
    prims<NDIM>(qL, WL, 0, (B)prm.gamma);
    prims<NDIM>(qR, WR, 0, (B)prm.gamma);
    riemann<NDIM, RS>(WL, WR, qL, qR, F, 0, (B)prm.gamma, (B)prm.min_c2, prm.lowmach_sd);
#pragma unroll
    for (int v = 0; v < V; ++v) buf[v][lp][fi] = F[v];
  }
  __syncthreads();
  // write: face grid of x, idx = (row) * (Na+1) + f with row = transverse coordinate
  long long plane = 1;
#pragma unroll
  for (int d = 0; d < NDIM; ++d) plane *= (d == 0) ? (long long)Na + 1 : (long long)g.N[d] * n;
  for (int k = threadIdx.x; k < V * NLP * FPB; k += 256) {
    const int v = k / (NLP * FPB), r = k % (NLP * FPB);
    const int l = r / FPB, fo = r % FPB;
    const int ff = fc * FPB + fo;
    if (ff > Na) continue;
    int cq[3] = {0, 0, 0};
    int rr = l;
#pragma unroll
    for (int d = 1; d < NDIM; ++d) {
      cq[d] = rr % n;
      rr /= n;
    }
    long long row = 0;
#pragma unroll
    for (int d = NDIM - 1; d >= 1; --d) row = row * ((long long)g.N[d] * n) + (long long)ce[d] * n + cq[d];
    Fface[(long long)(t * V + v) * plane + row * (Na + 1) + ff] = buf[v][l][fo];
  }
}

template <int NDIM, int P, int RS, int A, typename S, typename B>
__global__ void __launch_bounds__(256)
k_face_sd(Ops<P, B> op, Geom g, Params prm, const S* __restrict__ src, int Tsrc,
          B* __restrict__ Fface) {
  SPD_SKIP(g);
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
  const long long pidx = tid % plane;
  int ce[3] = {0, 0, 0}, cl[3] = {0, 0, 0};
  int f = 0;
  long long widx = pidx;  // face-grid index of the output
  if constexpr (NDIM == 3 && A == 2) {
    // z faces: the threads walk the points of an element's xy plane first, so
    // that the lanes of a wave read whole planes of the elements (its lines
    // along z are n*n values apart) instead of 4-value pieces of rows
    long long x = pidx;
    cl[0] = split_index(x, n);
    cl[1] = split_index(x, n);
    ce[0] = split_index(x, g.N[0]);
    ce[1] = split_index(x, g.N[1]);
    f = (int)x;
    widx = ((long long)f * ext[1] + (long long)ce[1] * n + cl[1]) * ext[0] + (long long)ce[0] * n +
           cl[0];
  } else {
    long long x = pidx;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      const int q = split_index(x, ext[b]);
      if (b == a) {
        f = q;
      } else {
        ce[b] = q / n;
        cl[b] = q % n;
      }
    }
  }
  int lb = 0;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) lb = lb * n + ((b == a) ? 0 : cl[b]);
  const int sa = stride_of<n>(a);
  const S* base = src + (long long)t * V * g.NE * NL;

  auto trace = [&](int ea, int fp, B* q) {
    int c[3] = {ce[0], ce[1], ce[2]};
    c[a] = ea;
    const long long e = elem_index<NDIM>(g, c);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const S* u = base + ((long long)v * g.NE + e) * NL + lb;
      B acc = B(0);
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * (B)u[s * sa];
      q[v] = acc;
    }
  };
  B qL[V], qR[V];
  if (f > 0) trace(f - 1, nf - 1, qL);
  if (f < Na) trace(f, 0, qR);
  if (f == 0) {
    const int k = g.bc[a][0];
    if (k == BC_PERIODIC) {
      trace(Na - 1, nf - 1, qL);
    } else if (bc_stored(k)) {
      bc_stored_traces<V, B>(g, a, 0, bc_point<NDIM>(g, a, n, ce, cl), qL, qR);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) qL[v] = qR[v];
      if (k == BC_REFLECTIVE) qL[a + 1] = -qL[a + 1];
    }
  }
  if (f == Na) {
    const int k = g.bc[a][1];
    if (k == BC_PERIODIC) {
      trace(0, 0, qR);
    } else if (bc_stored(k)) {
      bc_stored_traces<V, B>(g, a, 1, bc_point<NDIM>(g, a, n, ce, cl), qR, qL);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) qR[v] = qL[v];
      if (k == BC_REFLECTIVE) qR[a + 1] = -qR[a + 1];
    }
  }
  B WL[V], WR[V], F[V];
  prims<NDIM>(qL, WL, a, (B)prm.gamma);
  prims<NDIM>(qR, WR, a, (B)prm.gamma);
  riemann<NDIM, RS>(WL, WR, qL, qR, F, a, (B)prm.gamma, (B)prm.min_c2, prm.lowmach_sd);
#pragma unroll
  for (int v = 0; v < V; ++v) Fface[(long long)(t * V + v) * plane + widx] = F[v];
}

// ---------------------------------------------------------------------------
// 2-D element faces of both dimensions in one pass over the source (the
// faces of k_face_sd_x and k_face_sd<2, P, RS, 1> with the same values): a
// block takes a tile of EX x EY elements (FaceTile2) at one time node.  Phase 1: one
// thread per (variable, element) walks the element's rows once and forms the
// two x-end traces of every row and the two y-end traces of every column in
// LDS (every sum in the order of those kernels).  Phase 2: one thread per
// face point solves the x faces left of the tile's elements and the y faces
// below them, and the grid's last x / y face where the tile ends there.  A
// trace of an element outside the tile (the left / lower neighbour, a
// periodic wrap) is formed from the source.
// ---------------------------------------------------------------------------
template <int P> struct FaceTile2 {
  static constexpr int EX = 16, EY = (16 / (P + 1) > 0) ? 16 / (P + 1) : 1;
};

template <int P, int RS, typename S, typename B>
__device__ __forceinline__ void k_face_sd2d_body(const BlockIx bi, Ops<P, B> op, Geom g, Params prm, const S* __restrict__ src, B* __restrict__ Fx,
            B* __restrict__ Fy) {
  SPD_SKIP(g);
  constexpr int V = 4, n = P + 1, nf = P + 2, NL = n * n;
  constexpr int EX = FaceTile2<P>::EX, EY = FaceTile2<P>::EY;
  __shared__ B tr[4][V][EY][EX][n];  // end traces: x left, x right, y lower, y upper
  const int Nx = g.N[0], Ny = g.N[1];
  const int ntx = (Nx + EX - 1) / EX, nty = (Ny + EY - 1) / EY;
  // tiles in XCD column bands when the tile columns divide by 8 (blocks go
  // round-robin to the 8 XCDs: XCD k takes the columns [k W, (k + 1) W) of
  // every time node, row by row, so a tile's left and lower neighbours ran
  // on its XCD one and W blocks before it); otherwise launch order
  int t, txb, tyb;
  if (ntx % 8 == 0) {
    const int W = ntx >> 3, j = (int)(bi.x >> 3), jj = j % (W * nty);
    t = j / (W * nty);
    txb = (int)(bi.x & 7) * W + jj % W;
    tyb = jj / W;
  } else {
    t = (int)(bi.x / (unsigned)(ntx * nty));
    const int rb = (int)(bi.x % (unsigned)(ntx * nty));
    txb = rb % ntx;
    tyb = rb / ntx;
  }
  const int ex0 = txb * EX, ey0 = tyb * EY;
  const S* base = src + (long long)t * V * g.NE * NL;
  for (int it = threadIdx.x; it < V * EY * EX; it += 256) {
    const int exl = it % EX, r = it / EX, eyl = r % EY, v = r / EY;
    if (ex0 + exl >= Nx || ey0 + eyl >= Ny) continue;
    const S* u = base + ((long long)v * g.NE + (long long)(ey0 + eyl) * Nx + ex0 + exl) * NL;
    B yl[n], yu[n];
#pragma unroll
    for (int q = 0; q < n; ++q) yl[q] = yu[q] = B(0);
#pragma unroll
    for (int s = 0; s < n; ++s) {  // row s of the element
      B w[n];
#pragma unroll
      for (int q = 0; q < n; ++q) w[q] = (B)u[s * n + q];
      B xl = B(0), xr = B(0);
#pragma unroll
      for (int q = 0; q < n; ++q) {
        xl += op.s2f[0][q] * w[q];
        xr += op.s2f[nf - 1][q] * w[q];
      }
      tr[0][v][eyl][exl][s] = xl;
      tr[1][v][eyl][exl][s] = xr;
#pragma unroll
      for (int q = 0; q < n; ++q) {
        yl[q] += op.s2f[0][s] * w[q];
        yu[q] += op.s2f[nf - 1][s] * w[q];
      }
    }
#pragma unroll
    for (int q = 0; q < n; ++q) {
      tr[2][v][eyl][exl][q] = yl[q];
      tr[3][v][eyl][exl][q] = yu[q];
    }
  }
  __syncthreads();
  // trace of element (ex, ey) from the source: along x on row `line`
  // (axis 0) or along y on column `line` (axis 1), at flux point fp
  auto gtrace = [&](int ex, int ey, int fp, int axis, int line, B* q) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const S* u = base + ((long long)v * g.NE + (long long)ey * Nx + ex) * NL;
      B acc = B(0);
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * (B)u[axis == 0 ? line * n + s : s * n + line];
      q[v] = acc;
    }
  };
  const B gam = (B)prm.gamma, mc2 = (B)prm.min_c2;
  // x face f = ex0 + fl on row k of element row ey0 + eyl
  auto xface = [&](int fl, int k, int eyl) {
    const int f = ex0 + fl, ey = ey0 + eyl;
    if (ey >= Ny || f > Nx) return;
    B qL[V], qR[V];
    if (f > 0) {
      if (fl > 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) qL[v] = tr[1][v][eyl][fl - 1][k];
      } else {
        gtrace(f - 1, ey, nf - 1, 0, k, qL);
      }
    }
    if (f < Nx) {
#pragma unroll
      for (int v = 0; v < V; ++v) qR[v] = tr[0][v][eyl][fl][k];
    }
    if (f == 0) {
      const int b = g.bc[0][0];
      if (b == BC_PERIODIC) {
        gtrace(Nx - 1, ey, nf - 1, 0, k, qL);
      } else if (bc_stored(b)) {
        bc_stored_traces<V, B>(g, 0, 0, (long long)ey * n + k, qL, qR);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) qL[v] = qR[v];
        if (b == BC_REFLECTIVE) qL[1] = -qL[1];
      }
    }
    if (f == Nx) {
      const int b = g.bc[0][1];
      if (b == BC_PERIODIC) {
        gtrace(0, ey, 0, 0, k, qR);
      } else if (bc_stored(b)) {
        bc_stored_traces<V, B>(g, 0, 1, (long long)ey * n + k, qR, qL);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) qR[v] = qL[v];
        if (b == BC_REFLECTIVE) qR[1] = -qR[1];
      }
    }
    B WL[V], WR[V], F[V];
    prims<2>(qL, WL, 0, gam);
    prims<2>(qR, WR, 0, gam);
    riemann<2, RS>(WL, WR, qL, qR, F, 0, gam, mc2, prm.lowmach_sd);
    const long long plane = (long long)(Nx + 1) * Ny * n;
#pragma unroll
    for (int v = 0; v < V; ++v)
      Fx[(long long)(t * V + v) * plane + (long long)(ey * n + k) * (Nx + 1) + f] = F[v];
  };
  // y face f = ey0 + fl on column k of element ex0 + exl
  auto yface = [&](int fl, int exl, int k) {
    const int f = ey0 + fl, ex = ex0 + exl;
    if (ex >= Nx || f > Ny) return;
    B qL[V], qR[V];
    if (f > 0) {
      if (fl > 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) qL[v] = tr[3][v][fl - 1][exl][k];
      } else {
        gtrace(ex, f - 1, nf - 1, 1, k, qL);
      }
    }
    if (f < Ny) {
#pragma unroll
      for (int v = 0; v < V; ++v) qR[v] = tr[2][v][fl][exl][k];
    }
    if (f == 0) {
      const int b = g.bc[1][0];
      if (b == BC_PERIODIC) {
        gtrace(ex, Ny - 1, nf - 1, 1, k, qL);
      } else if (bc_stored(b)) {
        bc_stored_traces<V, B>(g, 1, 0, (long long)ex * n + k, qL, qR);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) qL[v] = qR[v];
        if (b == BC_REFLECTIVE) qL[2] = -qL[2];
      }
    }
    if (f == Ny) {
      const int b = g.bc[1][1];
      if (b == BC_PERIODIC) {
        gtrace(ex, 0, 0, 1, k, qR);
      } else if (bc_stored(b)) {
        bc_stored_traces<V, B>(g, 1, 1, (long long)ex * n + k, qR, qL);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v) qR[v] = qL[v];
        if (b == BC_REFLECTIVE) qR[2] = -qR[2];
      }
    }
    B WL[V], WR[V], F[V];
    prims<2>(qL, WL, 1, gam);
    prims<2>(qR, WR, 1, gam);
    riemann<2, RS>(WL, WR, qL, qR, F, 1, gam, mc2, prm.lowmach_sd);
    const long long plane = (long long)(Ny + 1) * Nx * n;
#pragma unroll
    for (int v = 0; v < V; ++v)
      Fy[(long long)(t * V + v) * plane + (long long)f * Nx * n + ex * n + k] = F[v];
  };
  // faces left of / below the tile's elements (face index fastest along x,
  // point index fastest along y: the lanes write along the face grids' rows)
  for (int it = threadIdx.x; it < EX * n * EY; it += 256) xface(it % EX, (it / EX) % n, it / (EX * n));
  if (ex0 + EX == Nx)
    for (int it = threadIdx.x; it < n * EY; it += 256) xface(EX, it % n, it / n);
  for (int it = threadIdx.x; it < EX * n * EY; it += 256) yface(it / (EX * n), (it % (EX * n)) / n, it % n);
  if (ey0 + EY == Ny)
    for (int it = threadIdx.x; it < EX * n; it += 256) yface(EY, it / n, it % n);
}
template <int P, int RS, typename S, typename B>
__global__ void __launch_bounds__(256)
k_face_sd2d(Ops<P, B> op, Geom g, Params prm, const S* __restrict__ src, B* __restrict__ Fx,
            B* __restrict__ Fy) {
  k_face_sd2d_body<P, RS, S, B>(hw_block(), op, g, prm, src, Fx, Fy);
}

// ---------------------------------------------------------------------------
// Predictor update.  A block owns EPB elements and all T time nodes.  Phase
// "dim a": one thread per (element, t, line along a) interpolates the line to
// the flux points, evaluates the Euler flux at the interior ones, takes the
// element-face Riemann fluxes at the ends and differentiates back
// (src/sdader_simulator.py:148-158, src/sd_simulator.py:190-206); the
// derivative accumulates in LDS.  Update phase:
//   mode 0 (Picard, src/sdader_simulator.py:187-200):
//      Uad[n] = U_sp - dt * sum_p invader[n,p] dudt[p]
//   mode 1 (SD corrector, src/sdader_simulator.py:202-208):
//      U_sp  -= dt * sum_t w_t dudt[t]
// With Tsrc == 1 the source is U_sp for every time node (first iteration)
// and only t = 0 is evaluated.  Mode 0 writes out_ad (bulk type), mode 1
// out_sp (FP64, accumulated in FP64).
// ---------------------------------------------------------------------------
template <int NDIM, int P, typename S, typename B>
__global__ void __launch_bounds__(256)
k_pred(Ops<P, B> op, Geom g, Params prm, const S* src, int Tsrc, const B* __restrict__ Ff0,
       const B* __restrict__ Ff1, const B* __restrict__ Ff2, const double* Usp,
       const double* __restrict__ dt_dev, B* out_ad, double* out_sp, int mode, int EPB) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2, T = P + 1;
  constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;
  // dudt [EPB][T] slabs of V*NL values, padded by one value so that lanes of
  // different time nodes / elements fall into different LDS banks
  constexpr int SLAB = V * NL + 1;
  B* lds = reinterpret_cast<B*>(spd_dyn_lds);
  const int el = threadIdx.x / (T * NLP);
  const int rem = threadIdx.x % (T * NLP);
  const int t = rem / NLP;
  const int lp = rem % NLP;
  const long long e = (long long)blockIdx.x * EPB + el;
  const bool active = (el < EPB) && (e < g.NE) && (t < Tsrc);
  const B* Ff[3] = {Ff0, Ff1, Ff2};
  const B gam = (B)prm.gamma;
  int c[3] = {0, 0, 0};
  if (e < g.NE) elem_coords<NDIM>(g, e, c);

#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    if (active) {
      const int lb = line_base<NDIM, n>(a, lp);
      const int sa = stride_of<n>(a);
      B u[V][n];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const S* s0 = src + ((long long)(t * V + v) * g.NE + e) * NL + lb;
#pragma unroll
        for (int s = 0; s < n; ++s) u[v][s] = (B)s0[s * sa];
      }
      B F[nf][V];
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
      long long plane;
      const long long i0 = face_grid_base<NDIM, n>(g, a, c, cl, c[a], &plane);
      const long long sfa = stride_face<NDIM, n>(g, a);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const B* fb = Ff[a] + (long long)(t * V + v) * plane + i0;
        F[0][v] = fb[0];
        F[nf - 1][v] = fb[sfa];
      }
#pragma unroll
      for (int fp = 1; fp < nf - 1; ++fp) {
        B M[V], W[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          B acc = B(0);
#pragma unroll
          for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * u[v][s];
          M[v] = acc;
        }
        prims<NDIM>(M, W, a, gam);
        flux<NDIM>(W, F[fp], a, gam);
      }
      const B rha = (B)g.rh[a];
      B* d0 = lds + (long long)(el * T + t) * SLAB + lb;
#pragma unroll
      for (int s = 0; s < n; ++s) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          B acc = B(0);
#pragma unroll
          for (int fp = 0; fp < nf; ++fp) acc += op.d2s[s][fp] * F[fp][v];
          const B d = acc * rha;
          B* dst = d0 + v * NL + s * sa;
          if (a == 0)
            *dst = d;
          else
            *dst = *dst + d;
        }
      }
    }
    __syncthreads();
  }

  const double dt = *dt_dev;
  const int total = EPB * V * NL;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int le = idx / (V * NL);
    const int v = (idx / NL) % V;
    const int l = idx % NL;
    const long long ee = (long long)blockIdx.x * EPB + le;
    if (ee >= g.NE) continue;
    const long long off = ((long long)v * g.NE + ee) * NL + l;
    const double u0 = Usp[off];
    const B* dd = lds + (long long)le * T * SLAB + v * NL + l;  // + p*SLAB
    if (mode == 0) {
#pragma unroll
      for (int tn = 0; tn < T; ++tn) {
        B acc = B(0);
#pragma unroll
        for (int p = 0; p < T; ++p) {
          const int ps = (p < Tsrc) ? p : (Tsrc - 1);
          acc += op.inv[tn][p] * dd[(long long)ps * SLAB];
        }
        out_ad[((long long)(tn * V + v) * g.NE + ee) * NL + l] = (B)(u0 - (double)acc * dt);
      }
    } else {
      double acc = 0.0;
#pragma unroll
      for (int p = 0; p < T; ++p) acc += op.w[p] * (double)dd[(long long)p * SLAB];
      out_sp[off] = u0 - acc * dt;
    }
  }
}

// ---------------------------------------------------------------------------
// High-order fluxes of time node t on the FV faces of dimension a: flux-point
// values (end points from the element-face Riemann solve), averaged over the
// subcell extent in the transverse dimensions (src/sdader_simulator.py:288-323,
// src/sd_simulator.py:208-212).  One thread per (element, flux point).
// ---------------------------------------------------------------------------
template <int NDIM, int P, int A, typename B>
__global__ void __launch_bounds__(256)
k_hoflux(Ops<P, B> op, Geom g, Params prm, int t, const B* __restrict__ Uad,
         const B* __restrict__ Fface, B* __restrict__ Fho) {
  SPD_SKIP(g);
  constexpr int a = A;
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.NE * nf) return;
  const int f = (int)(tid % nf);
  const long long e = tid / nf;
  int c[3] = {0, 0, 0};
  elem_coords<NDIM>(g, e, c);
  const int Na = g.N[a];
  if (f == nf - 1 && c[a] != Na - 1) return;
  const int sa = stride_of<n>(a);
  B Fl[NLP][V];
#pragma unroll
  for (int lp = 0; lp < NLP; ++lp) {
    if (f == 0 || f == nf - 1) {
      const int fa = c[a] + (f == 0 ? 0 : 1);
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
      long long plane;
      const long long i0 = face_grid_base<NDIM, n>(g, a, c, cl, fa, &plane);
#pragma unroll
      for (int v = 0; v < V; ++v) Fl[lp][v] = Fface[(long long)(t * V + v) * plane + i0];
    } else {
      const int lb = line_base<NDIM, n>(a, lp);
      B M[V], W[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const B* u = Uad + ((long long)(t * V + v) * g.NE + e) * NL + lb;
        B acc = B(0);
#pragma unroll
        for (int s = 0; s < n; ++s) acc += op.s2f[f][s] * u[s * sa];
        M[v] = acc;
      }
      prims<NDIM>(M, W, a, (B)prm.gamma);
      flux<NDIM>(W, Fl[lp], a, (B)prm.gamma);
    }
  }
  // transverse subcell averages, lower dimension first
  B Fa[NLP][V];
  if constexpr (NDIM == 1) {
#pragma unroll
    for (int v = 0; v < V; ++v) Fa[0][v] = Fl[0][v];
  } else if constexpr (NDIM == 2) {
#pragma unroll
    for (int q = 0; q < n; ++q)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        B acc = B(0);
#pragma unroll
        for (int s = 0; s < n; ++s) acc += op.s2c[q][s] * Fl[s][v];
        Fa[q][v] = acc;
      }
  } else {
    B G[NLP][V];
#pragma unroll
    for (int c2 = 0; c2 < n; ++c2)
#pragma unroll
      for (int q1 = 0; q1 < n; ++q1)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          B acc = B(0);
#pragma unroll
          for (int s = 0; s < n; ++s) acc += op.s2c[q1][s] * Fl[s + n * c2][v];
          G[q1 + n * c2][v] = acc;
        }
#pragma unroll
    for (int q2 = 0; q2 < n; ++q2)
#pragma unroll
      for (int q1 = 0; q1 < n; ++q1)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          B acc = B(0);
#pragma unroll
          for (int s = 0; s < n; ++s) acc += op.s2c[q2][s] * G[q1 + n * s][v];
          Fa[q1 + n * q2][v] = acc;
        }
  }
  // scatter into the FV face array of dimension a
  long long ext[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) ext[b] = (b < NDIM) ? (g.Ns[b] + (b == a ? 1 : 0)) : 1;
  long long nface = 1;
#pragma unroll
  for (int b = 0; b < NDIM; ++b) nface *= ext[b];
#pragma unroll
  for (int lp = 0; lp < NLP; ++lp) {
    int q[3] = {0, 0, 0};
    int r = lp;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      if (b == a) continue;
      q[b] = r % n;
      r /= n;
    }
    long long idx = 0;
#pragma unroll
    for (int b = NDIM - 1; b >= 0; --b) {
      const long long cb = (b == a) ? (long long)c[a] * n + f : (long long)c[b] * n + q[b];
      idx = idx * ext[b] + cb;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) Fho[v * nface + idx] = Fa[lp][v];
  }
}

// ---------------------------------------------------------------------------
// Element-layout tensor-product transform (sp <-> cv) of a (V, E, L) array,
// one thread per (element, variable, output point); the sum runs over all
// input points with the product of the 1-D factors (src/transforms.py:27-45).
// out_fv != 0 writes the FV layout instead of the element layout.
// ---------------------------------------------------------------------------
template <int NDIM, int P>
__global__ void __launch_bounds__(256)
k_tensor(Ops<P> op, Geom g, int to_cv, const double* __restrict__ in, double* __restrict__ out,
         int out_fv) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V, n = P + 1;
  constexpr int NL = IPow<n, NDIM>::v;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.NE * V * NL) return;
  const int lo = (int)(tid % NL);
  const long long r = tid / NL;
  const long long e = r % g.NE;
  const int v = (int)(r / g.NE);
  // to_cv: 1 sp -> cv, 0 cv -> sp, 2 sp -> regular subcell centres
  const double(*M)[n] = (to_cv == 1) ? op.s2c : ((to_cv == 2) ? op.s2r : op.c2s);
  int qo[3] = {0, 0, 0};
  {
    int x = lo;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      qo[b] = x % n;
      x /= n;
    }
  }
  const double* u = in + ((long long)v * g.NE + e) * NL;
  double acc = 0.0;
  for (int li = 0; li < NL; ++li) {
    int x = li;
    int qi[3];
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      qi[b] = x % n;
      x /= n;
    }
    double w = 1.0;
    // product in the order of the einsum operands (outermost dimension first)
#pragma unroll
    for (int b = NDIM - 1; b >= 0; --b) w = (b == NDIM - 1) ? M[qo[b]][qi[b]] : w * M[qo[b]][qi[b]];
    acc += w * u[li];
  }
  if (!out_fv) {
    out[((long long)v * g.NE + e) * NL + lo] = acc;
  } else {
    int c[3] = {0, 0, 0};
    elem_coords<NDIM>(g, e, c);
    int cc[3];
#pragma unroll
    for (int b = 0; b < NDIM; ++b) cc[b] = c[b] * n + qo[b];
    out[(long long)v * g.Nc + cell_index<NDIM>(g, cc)] = acc;
  }
}

// ---------------------------------------------------------------------------
// The same transforms for 1-D and 2-D with one thread per (variable,
// element): the element's values are loaded once (contiguous), transformed
// dimension by dimension in registers (the separable form of the tensor
// product; the summation order differs from the single sum above by
// round-off) and written once.  in_fv / out_fv select the FV layout for the
// input / output; in_copy (optional) receives the input in element layout.
// mode: 1 sp -> cv, 0 cv -> sp, 2 sp -> regular subcell centres.
// ---------------------------------------------------------------------------
template <int NDIM, int P>
__global__ void __launch_bounds__(256)
k_tensor_el(Ops<P> op, Geom g, int mode, const double* __restrict__ in, int in_fv,
            double* __restrict__ out, int out_fv, double* __restrict__ in_copy) {
  static_assert(NDIM <= 2, "1-D and 2-D");
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V, n = P + 1;
  constexpr int NY = (NDIM == 2) ? n : 1;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long long)V * g.NE) return;
  long long x = tid;
  const long long e = (g.NE <= 0x7fffffffll) ? split_index(x, (int)g.NE) : tid % g.NE;
  const int v = (g.NE <= 0x7fffffffll) ? (int)x : (int)(tid / g.NE);
  const double(*M)[n] = (mode == 1) ? op.s2c : ((mode == 2) ? op.s2r : op.c2s);
  long long ey = e;
  const int ex = split_index(ey, g.N[0]);
  // point (j, i) of the element: element layout e*NL + j*n + i; FV layout
  // (ey*n + j) * Ns_x + ex*n + i
  auto idx = [&](int fv, int j, int i) -> long long {
    if (fv) return (long long)v * g.Nc + (ey * NY + j) * g.Ns[0] + (long long)ex * n + i;
    return ((long long)v * g.NE + e) * (NY * n) + j * n + i;
  };
  double u[NY][n];
#pragma unroll
  for (int j = 0; j < NY; ++j)
#pragma unroll
    for (int i = 0; i < n; ++i) u[j][i] = in[idx(in_fv, j, i)];
  if (in_copy) {
#pragma unroll
    for (int j = 0; j < NY; ++j)
#pragma unroll
      for (int i = 0; i < n; ++i) in_copy[idx(0, j, i)] = u[j][i];
  }
  double t[NY][n];
#pragma unroll
  for (int j = 0; j < NY; ++j)
#pragma unroll
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < n; ++k) acc += M[i][k] * u[j][k];
      t[j][i] = acc;
    }
#pragma unroll
  for (int jo = 0; jo < NY; ++jo)
#pragma unroll
    for (int i = 0; i < n; ++i) {
      double acc;
      if constexpr (NDIM == 2) {
        acc = 0.0;
#pragma unroll
        for (int k = 0; k < n; ++k) acc += M[jo][k] * t[k][i];
      } else {
        acc = t[jo][i];
      }
      out[idx(out_fv, jo, i)] = acc;
    }
}

// Maximum wave speed of the new subcell-mean state (src/sd_simulator.py:
// 214-227) per block, and the optional W_cv snapshot in element layout; one
// thread per cell of the FV layout.
template <int NDIM>
__global__ void __launch_bounds__(256)
k_wave_cells(Geom g, Params prm, int n, const double* __restrict__ Ucv,
             double* __restrict__ Wcv_sd, double* __restrict__ partial) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V, P_ = Dims<NDIM>::P;
  const long long ci = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double cw = 0.0;
  if (ci < g.Nc) {
    double u[V], W[V];
#pragma unroll
    for (int v = 0; v < V; ++v) u[v] = Ucv[v * g.Nc + ci];
    prims<NDIM>(u, W, 0, prm.gamma);
    if (Wcv_sd) {
      int c[3] = {0, 0, 0};
      long long x = ci;
#pragma unroll
      for (int a = 0; a < NDIM; ++a) {
        c[a] = split_index(x, g.Ns[a]);
      }
      long long e = 0, l = 0;
#pragma unroll
      for (int a = NDIM - 1; a >= 0; --a) {
        e = e * g.N[a] + c[a] / n;
        l = l * n + c[a] % n;
      }
      long long NL = 1;
#pragma unroll
      for (int a = 0; a < NDIM; ++a) NL *= n;
#pragma unroll
      for (int v = 0; v < V; ++v) Wcv_sd[((long long)v * g.NE + e) * NL + l] = W[v];
    }
    cw = sound(W[P_], W[0], prm.gamma, prm.min_c2) * NDIM;
#pragma unroll
    for (int k = 1; k <= NDIM; ++k) cw = cw + fabs(W[k]);
  }
  cw = block_max(cw);
  if (threadIdx.x == 0) partial[blockIdx.x] = cw;
}

// ---------------------------------------------------------------------------
// 3-D high-order fluxes of one dimension (the arithmetic of k_hoflux, same
// order): n adjacent lanes per (element, flux point), lane c2 takes the n lines
// whose second transverse coordinate is c2, averages them over the first,
// and the n threads of a group exchange those partial averages (shuffles
// among adjacent lanes) for the average over the second; thread c2 writes the
// faces of second coordinate c2.  A wave holds ~3 elements instead of ~13:
// their slabs stay in L1 for lines across the element (z), and a thread holds
// n lines instead of n^2.
// ---------------------------------------------------------------------------
template <int P, int A, typename B>
__global__ void __launch_bounds__(256)
k_hoflux3(Ops<P, B> op, Geom g, Params prm, int t, const B* __restrict__ Uad,
          const B* __restrict__ Fface, B* __restrict__ Fho) {
  SPD_SKIP(g);
  constexpr int NDIM = 3, a = A;
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v;
  // groups of n adjacent lanes, GPW per wave (a group never straddles waves)
  constexpr int GPW = 64 / n;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)(tid % 64);
  if (lane >= GPW * n) return;
  const long long grp = (tid / 64) * GPW + lane / n;
  if (grp >= g.NE * nf) return;  // the whole group
  const int c2 = lane % n;
  const int f = (int)(grp % nf);
  const long long e = grp / nf;
  int c[3] = {0, 0, 0};
  elem_coords<NDIM>(g, e, c);
  const int Na = g.N[a];
  if (f == nf - 1 && c[a] != Na - 1) return;  // the whole group
  const int sa = stride_of<n>(a);
  B Fl[n][V];
#pragma unroll
  for (int s1 = 0; s1 < n; ++s1) {
    const int lp = s1 + n * c2;
    if (f == 0 || f == nf - 1) {
      const int fa = c[a] + (f == 0 ? 0 : 1);
      int cl[3] = {0, 0, 0};
      {
        int rr = lp;
#pragma unroll
        for (int b = 0; b < NDIM; ++b) {
          if (b == a) continue;
          cl[b] = rr % n;
          rr /= n;
        }
      }
      long long plane;
      const long long i0 = face_grid_base<NDIM, n>(g, a, c, cl, fa, &plane);
#pragma unroll
      for (int v = 0; v < V; ++v) Fl[s1][v] = Fface[(long long)(t * V + v) * plane + i0];
    } else {
      const int lb = line_base<NDIM, n>(a, lp);
      B M[V], W[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const B* u = Uad + ((long long)(t * V + v) * g.NE + e) * NL + lb;
        B acc = B(0);
#pragma unroll
        for (int s = 0; s < n; ++s) acc += op.s2f[f][s] * u[s * sa];
        M[v] = acc;
      }
      prims<NDIM>(M, W, a, (B)prm.gamma);
      flux<NDIM>(W, Fl[s1], a, (B)prm.gamma);
    }
  }
  // average over the first transverse coordinate (G of k_hoflux, column c2)
  B G[n][V];
#pragma unroll
  for (int q1 = 0; q1 < n; ++q1)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      B acc = B(0);
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.s2c[q1][s] * Fl[s][v];
      G[q1][v] = acc;
    }
  // over the second: Fa[q1 + n q2] with q2 = c2, from the group's columns
  const int lane0 = lane - c2;  // the group's first lane
  B Fa[n][V];
#pragma unroll
  for (int q1 = 0; q1 < n; ++q1)
#pragma unroll
    for (int v = 0; v < V; ++v) Fa[q1][v] = B(0);
#pragma unroll
  for (int s = 0; s < n; ++s) {
    const B w = op.s2c[c2][s];
#pragma unroll
    for (int q1 = 0; q1 < n; ++q1)
#pragma unroll
      for (int v = 0; v < V; ++v) Fa[q1][v] += w * __shfl(G[q1][v], lane0 + s);
  }
  long long ext[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) ext[b] = g.Ns[b] + (b == a ? 1 : 0);
  const long long nface = ext[0] * ext[1] * ext[2];
#pragma unroll
  for (int q1 = 0; q1 < n; ++q1) {
    const int lp = q1 + n * c2;
    int q[3] = {0, 0, 0};
    int rr = lp;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      if (b == a) continue;
      q[b] = rr % n;
      rr /= n;
    }
    long long idx = 0;
#pragma unroll
    for (int b = NDIM - 1; b >= 0; --b) {
      const long long cb = (b == a) ? (long long)c[a] * n + f : (long long)c[b] * n + q[b];
      idx = idx * ext[b] + cb;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) Fho[v * nface + idx] = Fa[q1][v];
  }
}

// ---------------------------------------------------------------------------
// 2-D sp <-> cv transforms of the time step through LDS: a block owns a row
// segment of T2D_EB(P) elements (about 256 cells), loads all variables of its
// cells with coalesced reads in the input layout, transforms them with the
// separable passes of k_tensor_el (same arithmetic, same order) and writes
// the output layout with coalesced stores.  cv -> sp (mode 0) also evaluates
// the next dt's wave speed of every cell from the loaded FV state (as
// k_wave_cells: block maxima in `partial`) and the optional snapshots (the
// input in element layout, its primitives in element layout).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int t2d_eb(int P) {
  return 256 / ((P + 1) * (P + 1)) > 0 ? 256 / ((P + 1) * (P + 1)) : 1;
}
inline long long t2d_blocks(const Geom& g, int P) {
  const int eb = t2d_eb(P);
  return (long long)g.N[1] * ((g.N[0] + eb - 1) / eb);
}

template <int P> struct T2D {
  static constexpr int V = 4, n = P + 1, NL = n * n, EB = t2d_eb(P), WP = EB * n + 1;
  using Tile = double[V][n][WP];  // [variable][row j][x], x = element * n + i
};

// One row segment; FULL: ne == EB (constant index arithmetic).
template <int P, int MODE, bool FULL>
__device__ __forceinline__ void tensor2d_segment(const Ops<P>& op, const Geom& g, const Params& prm,
                                                 const double* __restrict__ in,
                                                 double* __restrict__ out,
                                                 double* __restrict__ in_copy,
                                                 double* __restrict__ Wcv_sd,
                                                 double* __restrict__ partial, int ey, int ex0,
                                                 int ne_, typename T2D<P>::Tile& su,
                                                 typename T2D<P>::Tile& st) {
  using T = T2D<P>;
  constexpr int V = T::V, n = T::n, NL = T::NL, EB = T::EB;
  const int ne = FULL ? EB : ne_;
  const int wc = ne * n;  // cells of the segment along x
  const long long e0 = (long long)ey * g.N[0] + ex0;
  const long long row0 = (long long)ey * n * g.Ns[0] + (long long)ex0 * n;  // FV index of (j=0, x=0)
  // every load of the thread is issued before the first is used (256 threads)
  constexpr int NIT = (V * n * EB * n + 255) / 256;
  double ld[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int k = threadIdx.x + it * 256;
    ld[it] = 0.0;
    if (k < V * n * wc) {
      const int v = k / (n * wc), r = k % (n * wc);
      if (MODE == 0)  // FV layout: rows of wc cells
        ld[it] = in[(long long)v * g.Nc + row0 + (long long)(r / wc) * g.Ns[0] + r % wc];
      else  // element layout: ne * NL contiguous values per variable
        ld[it] = in[((long long)v * g.NE + e0) * NL + r];
    }
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int k = threadIdx.x + it * 256;
    if (k < V * n * wc) {
      const int v = k / (n * wc), r = k % (n * wc);
      if (MODE == 0)
        su[v][r / wc][r % wc] = ld[it];
      else
        su[v][(r % NL) / n][(r / NL) * n + (r % NL) % n] = ld[it];
    }
  }
  __syncthreads();
  const double(*M)[n] = (MODE == 1) ? op.s2c : op.c2s;
  for (int k = threadIdx.x; k < V * n * ne; k += 256) {  // x pass
    const int v = k / (n * ne), r = k % (n * ne), j = r / ne, el = r % ne;
    double u[n];
#pragma unroll
    for (int i = 0; i < n; ++i) u[i] = su[v][j][el * n + i];
#pragma unroll
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[i][q] * u[q];
      st[v][j][el * n + i] = acc;
    }
  }
  if (MODE == 0) {
    double cw = 0.0;
    for (int k = threadIdx.x; k < n * wc; k += 256) {
      const int j = k / wc, x = k % wc;
      double u[V], W[V];
#pragma unroll
      for (int v = 0; v < V; ++v) u[v] = su[v][j][x];
      prims<2>(u, W, 0, prm.gamma);
      if (Wcv_sd) {
        const long long e = e0 + x / n;
        const int l = j * n + x % n;
#pragma unroll
        for (int v = 0; v < V; ++v) Wcv_sd[((long long)v * g.NE + e) * NL + l] = W[v];
      }
      double c = sound(W[3], W[0], prm.gamma, prm.min_c2) * 2;
      c = c + fabs(W[1]);
      c = c + fabs(W[2]);
      cw = nmax(cw, c);
    }
    cw = block_max(cw);
    if (threadIdx.x == 0) *partial = cw;  // the block's own slot
    if (in_copy) {
      for (int k = threadIdx.x; k < V * n * wc; k += 256) {
        const int v = k / (n * wc), r = k % (n * wc), el = r / NL, l = r % NL;
        in_copy[((long long)v * g.NE + e0) * NL + r] = su[v][l / n][el * n + l % n];
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < V * wc; k += 256) {  // y pass, one column per item
    const int v = k / wc, x = k % wc;
    double t[n];
#pragma unroll
    for (int q = 0; q < n; ++q) t[q] = st[v][q][x];
#pragma unroll
    for (int jo = 0; jo < n; ++jo) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[jo][q] * t[q];
      st[v][jo][x] = acc;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < V * n * wc; k += 256) {
    const int v = k / (n * wc), r = k % (n * wc);
    if (MODE == 1)
      out[(long long)v * g.Nc + row0 + (long long)(r / wc) * g.Ns[0] + r % wc] = st[v][r / wc][r % wc];
    else
      out[((long long)v * g.NE + e0) * NL + r] = st[v][(r % NL) / n][(r / NL) * n + (r % NL) % n];
  }
}

// MODE: 1 sp -> cv, 0 cv -> sp (with the wave speed).  Blocks of 256 threads.
template <int P, int MODE>
__device__ __forceinline__ void k_tensor2d_body(const BlockIx bi, Ops<P> op, Geom g, Params prm, const double* __restrict__ in, double* __restrict__ out,
           double* __restrict__ in_copy, double* __restrict__ Wcv_sd, double* __restrict__ partial) {
  SPD_SKIP(g);
  constexpr int EB = T2D<P>::EB;
  __shared__ typename T2D<P>::Tile su, st;  // input; after the x pass, then the output
  const int nbx = (g.N[0] + EB - 1) / EB;
  const int ey = (int)(bi.x / nbx), ex0 = (int)(bi.x % nbx) * EB;
  const int ne = (g.N[0] - ex0 < EB) ? g.N[0] - ex0 : EB;
  double* pblk = partial ? partial + bi.x : partial;  // the block's maximum (cv -> sp)
  if (ne == EB)
    tensor2d_segment<P, MODE, true>(op, g, prm, in, out, in_copy, Wcv_sd, pblk, ey, ex0, ne, su,
                                    st);
  else
    tensor2d_segment<P, MODE, false>(op, g, prm, in, out, in_copy, Wcv_sd, pblk, ey, ex0, ne,
                                     su, st);
}
template <int P, int MODE>
__global__ void __launch_bounds__(256)
k_tensor2d(Ops<P> op, Geom g, Params prm, const double* __restrict__ in, double* __restrict__ out,
           double* __restrict__ in_copy, double* __restrict__ Wcv_sd, double* __restrict__ partial) {
  k_tensor2d_body<P, MODE>(hw_block(), op, g, prm, in, out, in_copy, Wcv_sd, partial);
}

// ---------------------------------------------------------------------------
// 3-D sp <-> cv transforms of the time step: a block owns t3d_eb(P)
// consecutive elements (about 256 points), stages all their variables in LDS
// and applies the 1-D operator along x, then y, then z (one thread per point,
// each pass through LDS; the separable form of k_tensor's single sum, equal
// to it up to round-off).  MODE 1 sp -> cv (element layout in, FV layout out);
// MODE 0 cv -> sp with the next dt's wave speed of every cell (block maxima in
// `partial`, as k_wave_cells) and the optional snapshots (the input and its
// primitives in element layout).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int t3d_eb(int P) {
  return 256 / ((P + 1) * (P + 1) * (P + 1)) > 0 ? 256 / ((P + 1) * (P + 1) * (P + 1)) : 1;
}

template <int P, int MODE>
__global__ void __launch_bounds__(256)
k_tensor3d(Ops<P> op, Geom g, Params prm, const double* __restrict__ in, double* __restrict__ out,
           double* __restrict__ in_copy, double* __restrict__ Wcv_sd, double* __restrict__ partial) {
  SPD_SKIP(g);
  constexpr int V = 5, n = P + 1, NL = n * n * n, EB = t3d_eb(P), NP = EB * NL;
  __shared__ double sa[V][NP], sb[V][NP];
  const long long e0 = (long long)blockIdx.x * EB;
  const int ne = (g.NE - e0 < EB) ? (int)(g.NE - e0) : EB;
  // cell index of point l of block element el (FV layout)
  auto cell_of = [&](int el, int l) -> long long {
    int ce[3];
    elem_coords<3>(g, e0 + el, ce);
    const int cc[3] = {ce[0] * n + l % n, ce[1] * n + (l / n) % n, ce[2] * n + l / (n * n)};
    return cell_index<3>(g, cc);
  };
  for (int k = threadIdx.x; k < V * ne * NL; k += blockDim.x) {
    const int v = k / (ne * NL), r = k % (ne * NL);
    sa[v][r] = (MODE == 1) ? in[((long long)v * g.NE + e0) * NL + r]
                           : in[(long long)v * g.Nc + cell_of(r / NL, r % NL)];
  }
  __syncthreads();
  const double(*M)[n] = (MODE == 1) ? op.s2c : op.c2s;
  const int k = threadIdx.x;
  const bool act = k < ne * NL;
  const int el = k / NL, l = k % NL;
  const int i = l % n, j = (l / n) % n, z = l / (n * n);
  const int b0 = el * NL;
  double cw = 0.0;
  if (MODE == 0 && act) {
    double u[V], W[V];
#pragma unroll
    for (int v = 0; v < V; ++v) u[v] = sa[v][k];
    prims<3>(u, W, 0, prm.gamma);
    const long long o = (e0 + el) * NL + l;
    if (in_copy) {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        in_copy[(long long)v * g.NE * NL + o] = u[v];
        Wcv_sd[(long long)v * g.NE * NL + o] = W[v];
      }
    }
    cw = sound(W[4], W[0], prm.gamma, prm.min_c2) * 3;
#pragma unroll
    for (int q = 1; q <= 3; ++q) cw = cw + fabs(W[q]);
  }
  if (act) {  // x
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[i][q] * sa[v][b0 + q + n * (j + n * z)];
      sb[v][k] = acc;
    }
  }
  __syncthreads();
  if (act) {  // y
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[j][q] * sb[v][b0 + i + n * (q + n * z)];
      sa[v][k] = acc;
    }
  }
  __syncthreads();
  if (act) {  // z, written out
    const long long ci = (MODE == 1) ? cell_of(el, l) : 0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < n; ++q) acc += M[z][q] * sa[v][b0 + i + n * (j + n * q)];
      if (MODE == 1)
        out[(long long)v * g.Nc + ci] = acc;
      else
        out[((long long)v * g.NE + e0) * NL + k] = acc;
    }
  }
  if (MODE == 0) {
    cw = block_max(cw);
    if (threadIdx.x == 0) partial[blockIdx.x] = cw;
  }
}

// ---------------------------------------------------------------------------
// FV update U_out = U_in - dt_i * sum_a (F_a[f+1] - F_a[f]) / dx_a
// (src/fv_simulator.py:189-205); in place when U_out == U_in.
// ---------------------------------------------------------------------------
template <int NDIM, typename B>
__global__ void __launch_bounds__(256)
k_apply(Geom g, const double* Uin, const B* __restrict__ F0, const B* __restrict__ F1,
        const B* __restrict__ F2, const double* __restrict__ dt_dev, double w_i, double* Uout) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.Nc) return;
  int c[3] = {0, 0, 0};
  {
    long long x = cell;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      c[a] = split_index(x, g.Ns[a]);
    }
  }
  const B* F[3] = {F0, F1, F2};
  long long fi[3], fs[3], nf[3];
  double dx[3];
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    long long idx = 0, stride = 1, sa = 0, tot = 1;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      const long long ext = g.Ns[b] + (b == a ? 1 : 0);
      if (b == a) sa = stride;
      idx += c[b] * stride;
      stride *= ext;
      tot *= ext;
    }
    fi[a] = idx;
    fs[a] = sa;
    nf[a] = tot;
    dx[a] = g.rhfp[a][c[a] + NGH];  // reciprocal width
  }
  const double dti = (*dt_dev) * w_i;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const double u = Uin[v * g.Nc + cell];
    double d = u * 0.0;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      const B* Fa = F[a] + v * nf[a] + fi[a];
      d = d + ((double)Fa[fs[a]] - (double)Fa[0]) * dx[a];
    }
    Uout[v * g.Nc + cell] = u - d * dti;
  }
}

// ---------------------------------------------------------------------------
// Trouble detection per active subcell (src/trouble_detection.py:5-71):
// NAD against the plus-stencil extrema of the ghost-filled old primitives,
// smooth-extrema relaxation (p > 1), PAD bounds.  Sets *any when any variable
// of any cell is a NAD candidate (the reference's global gate, Q3).
// ---------------------------------------------------------------------------
// Primitives of the FV state entering a substep (V x Nc), one thread per
// active cell: detection and the Hancock states read them on their stencils.
template <int NDIM>
__global__ void __launch_bounds__(256)
k_prims_cells(Geom g, Params prm, const double* __restrict__ U, double* __restrict__ W) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.Nc) return;
  double u[V], w[V];
#pragma unroll
  for (int v = 0; v < V; ++v) u[v] = U[v * g.Nc + cell];
  prims<NDIM>(u, w, 0, prm.gamma);
#pragma unroll
  for (int v = 0; v < V; ++v) W[v * g.Nc + cell] = w[v];
}

// Density and pressure of the candidate state, one thread per active cell:
// the smooth-extrema test reads them up to three cells away along every
// dimension (rp: 2 x Nc).
template <int NDIM>
__global__ void __launch_bounds__(256)
k_rho_p(Geom g, Params prm, const double* __restrict__ Unew, double* __restrict__ rp) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V, P = Dims<NDIM>::P;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= g.Nc) return;
  double u[V], W[V];
#pragma unroll
  for (int v = 0; v < V; ++v) u[v] = Unew[v * g.Nc + cell];
  prims<NDIM>(u, W, 0, prm.gamma);
  rp[cell] = W[0];
  rp[g.Nc + cell] = W[P];
}

// alpha of the smooth-extrema relaxation of variable k (0: density, 1:
// pressure, read from rp) at the active cell c along dimension a.
template <int NDIM>
__device__ __forceinline__ double sed_alpha(const double* __restrict__ rp, const Geom& g,
                                            const int* c, int a, int k) {
  const int Ns = g.Ns[a];
  const int ca = c[a];
  const double* rdc = g.rdc[a] + NGH;  // 1 / (centre spacing over two cells), active-relative
  const double* hf = g.hfp[a] + NGH;
  const double* r = rp + k * g.Nc;
  double w[7];
#pragma unroll
  for (int k2 = -3; k2 <= 3; ++k2) {
    const int x = ca + k2;
    if (x >= -NGH && x < Ns + NGH) {
      int cc[3] = {c[0], c[1], c[2]};
      bool fl = false;
      if ((x < 0 || x >= Ns) && bc_stored(g.bc[a][x < 0 ? 0 : 1])) {
        cc[a] = x;  // the stored layers' density / pressure
        w[k2 + 3] = g.wgh[(k == 0 ? 0 : NDIM + 1) * g.Ngc + ghosted_index<NDIM>(g, cc)];
        continue;
      }
      cc[a] = ghost_map(x, Ns, g.bc[a], &fl);  // rho and P are even under the flips
      w[k2 + 3] = r[cell_index<NDIM>(g, cc)];
    } else {
      w[k2 + 3] = 0.0;
    }
  }
  double dU[5];  // relative -2..2
#pragma unroll
  for (int k = -2; k <= 2; ++k) {
    const int x = ca + k;
    if (x >= -1 && x <= Ns)
      dU[k + 2] = (w[k + 4] - w[k + 2]) * rdc[x];
    else
      dU[k + 2] = 0.0;
  }
  double A[3], aLc = 1.0;
#pragma unroll
  for (int k = -1; k <= 1; ++k) {
    const int x = ca + k;
    if (x < 0 || x >= Ns) {
      A[k + 1] = 0.0;
      continue;
    }
    const double d2U = (dU[k + 3] - dU[k + 1]) * rdc[x];
    const double dv = 0.5 * hf[x] * d2U;
    const double vL = dU[k + 1] - dU[k + 2];
    const double vR = dU[k + 3] - dU[k + 2];
    double aL = -((dv < 0.0) ? ((vL > 0.0) ? vL : 0.0) : ((vL < 0.0) ? vL : 0.0)) / dv;
    aL = (fabs(dv) <= 0.0) ? 1.0 : aL;
    aL = (aL < 1.0) ? aL : 1.0;
    double aR = ((dv > 0.0) ? ((vR > 0.0) ? vR : 0.0) : ((vR < 0.0) ? vR : 0.0)) / dv;
    aR = (fabs(dv) <= 0.0) ? 1.0 : aR;
    aR = (aR < 1.0) ? aR : 1.0;
    A[k + 1] = (aR < aL) ? aR : aL;
    if (k == 0) aLc = aL;
  }
  // compute_min with the reference's edge rules (Q5)
  if (Ns == 1) return aLc;
  if (ca == 0) return nmin(A[1], A[2]);
  if (ca == Ns - 1) return nmin(A[0], aLc);
  return nmin(A[0], nmin(A[1], A[2]));
}

template <int NDIM>
__global__ void __launch_bounds__(256)
k_detect(Geom g, Params prm, const double* __restrict__ Wold, const double* __restrict__ Unew,
         const double* __restrict__ rp, uint8_t* __restrict__ flags, int* __restrict__ any_flag) {
  SPD_SKIP(g);
  constexpr int V = Dims<NDIM>::V, P = Dims<NDIM>::P;
  const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool cand = false;
  if (cell < g.Nc) {
    int c[3] = {0, 0, 0};
    {
      long long x = cell;
#pragma unroll
      for (int a = 0; a < NDIM; ++a) {
        c[a] = split_index(x, g.Ns[a]);
      }
    }
    double Wn[V], Wo[V], mx[V], mn[V];
    load_W<NDIM>(Unew, g, c, Wn, prm.gamma);
    load_prim<NDIM>(Wold, g, c, Wo);
#pragma unroll
    for (int v = 0; v < V; ++v) mx[v] = mn[v] = Wo[v];
#pragma unroll
    for (int a = 0; a < NDIM; ++a)
#pragma unroll
      for (int s = -1; s <= 1; s += 2) {
        int cc[3] = {c[0], c[1], c[2]};
        cc[a] += s;
        double Wb[V];
        load_prim<NDIM>(Wold, g, cc, Wb);
#pragma unroll
        for (int v = 0; v < V; ++v) {
          mx[v] = nmax(mx[v], Wb[v]);
          mn[v] = nmin(mn[v], Wb[v]);
        }
      }
    bool pt[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double lo = mn[v], hi = mx[v];
      if (prm.p > 0) {
        if (prm.nad_delta) {
          const double eps = prm.tol * (hi - lo);
          lo = lo - eps;
          hi = hi + eps;
        } else {
          lo = lo - fabs(lo) * prm.tol;
          hi = hi + fabs(hi) * prm.tol;
        }
      }
      pt[v] = !(Wn[v] >= lo) || !(Wn[v] <= hi);
      cand = cand || pt[v];
    }
    if (prm.p > 1 && prm.sed) {
      const int vars[2] = {0, P};
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int var = vars[k];
        if (!pt[var]) continue;
        double alpha = 1.0;
#pragma unroll
        for (int a = 0; a < NDIM; ++a) {
          const double an = sed_alpha<NDIM>(rp, g, c, a, k);
          alpha = (an < alpha) ? an : alpha;
        }
        pt[var] = pt[var] && (alpha < 1.0);
      }
    }
    bool tr = pt[0] || pt[P];
    if (prm.pad) {
      if (!(Wn[0] >= prm.min_rho) || !(Wn[0] <= prm.max_rho) || !(Wn[P] >= prm.min_P)) tr = true;
    }
    flags[cell] = tr ? 1 : 0;
  }
  // the global gate only matters when some side is not periodic
  if (any_flag) {
    const int blk = __syncthreads_or(cand ? 1 : 0);
    if (blk && threadIdx.x == 0 && *(volatile int*)any_flag == 0) atomicOr(any_flag, 1);
  }
}

// Trouble value on the ghosted grid after the flag fill that refreshes only
// periodic sides (src/trouble_detection.py:73-76, src/fv_simulator.py:224-273,
// Q2): a ghost of a non-periodic side keeps the density of the last full
// primitive fill (the new state if the smooth-extrema branch ran, else the old).
template <int NDIM>
__device__ __forceinline__ double trouble_at(const Geom& g, const uint8_t* __restrict__ flags,
                                             const double* __restrict__ Ulast, const int* cg) {
  int m[3] = {cg[0], cg[1], cg[2]};
#pragma unroll
  for (int a = NDIM - 1; a >= 0; --a) {
    const int Ns = g.Ns[a];
    if (m[a] < 0 || m[a] >= Ns) {
      const int side = (m[a] < 0) ? 0 : 1;
      if (bc_stored(g.bc[a][side]))  // the stored layers' density
        return g.wgh[ghosted_index<NDIM>(g, m)];
      if (g.bc[a][side] != BC_PERIODIC) {
        m[a] = (side == 0) ? m[a] + NGH : m[a] - NGH;
#pragma unroll
        for (int b = a - 1; b >= 0; --b) {
          // the stale layers were copied from a state whose lower-dimension
          // ghosts a stored side had filled with its own layers
          if ((m[b] < 0 || m[b] >= g.Ns[b]) && bc_stored(g.bc[b][m[b] < 0 ? 0 : 1]))
            return g.wgh[ghosted_index<NDIM>(g, m)];
          bool fl = false;
          m[b] = ghost_map(m[b], g.Ns[b], g.bc[b], &fl);
        }
        return Ulast[cell_index<NDIM>(g, m)];
      }
      m[a] = (side == 0) ? m[a] + Ns : m[a] - Ns;
    }
  }
  return (double)flags[cell_index<NDIM>(g, m)];
}

// Blending stencil: (target slice per dim: 0 full, 1 = [:-1], 2 = [1:]),
// source offset per dim, weight (src/trouble_detection.py:139-169, Q6).
// Coordinates of linear index x on a grid of extents ext[0..NDIM) (x
// fastest); 32-bit arithmetic when the grid has fewer than 2^32 points.
template <int NDIM>
__device__ __forceinline__ void grid_coords(long long x, const int* ext, long long total, int* c) {
  if (total <= 0xffffffffll) {
    unsigned u = (unsigned)x;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      c[a] = (int)(u % (unsigned)ext[a]);
      u /= (unsigned)ext[a];
    }
  } else {
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      c[a] = split_index(x, ext[a]);
    }
  }
}

struct BlendRule {
  signed char tgt[3];
  signed char off[3];
  double w;
};
// 2-D diagonals, dims ordered (x, y)
__constant__ BlendRule kRules2[4] = {
    {{1, 1, 0}, {+1, +1, 0}, 0.5},
    {{2, 1, 0}, {-1, +1, 0}, 0.5},
    {{1, 2, 0}, {+1, -1, 0}, 0.5},
    {{2, 2, 0}, {-1, -1, 0}, 0.5},
};
// 3-D: per plane (y,x), (z,y), (x,z) the four updates of the reference's loop
// (two diagonals, two axis neighbours), then the four corners with dz = +1.
__constant__ BlendRule kRules3[16] = {
    // plane idim=0: A=y, B=x
    {{1, 1, 0}, {+1, +1, 0}, 0.5},
    {{2, 1, 0}, {-1, 0, 0}, 0.5},
    {{1, 2, 0}, {+1, -1, 0}, 0.5},
    {{2, 2, 0}, {-1, 0, 0}, 0.5},
    // plane idim=1: A=z, B=y
    {{0, 1, 1}, {0, +1, +1}, 0.5},
    {{0, 2, 1}, {0, -1, 0}, 0.5},
    {{0, 1, 2}, {0, +1, -1}, 0.5},
    {{0, 2, 2}, {0, -1, 0}, 0.5},
    // plane idim=2: A=x, B=z
    {{1, 0, 1}, {+1, 0, +1}, 0.5},
    {{1, 0, 2}, {0, 0, -1}, 0.5},
    {{2, 0, 1}, {-1, 0, +1}, 0.5},
    {{2, 0, 2}, {0, 0, -1}, 0.5},
    // corners
    {{1, 1, 1}, {+1, +1, +1}, 0.375},
    {{2, 1, 1}, {-1, +1, +1}, 0.375},
    {{1, 2, 1}, {+1, -1, +1}, 0.375},
    {{2, 2, 1}, {-1, -1, +1}, 0.375},
};

__device__ __forceinline__ bool in_slice(int gx, int Ng, int kind) {
  return kind == 0 || (kind == 1 && gx <= Ng - 2) || (kind == 2 && gx >= 1);
}

// theta before the last ring: trouble maximum-blended with its neighbours;
// without blending, theta = trouble (and k_ring is skipped).
template <int NDIM>
__global__ void __launch_bounds__(256)
k_theta1(Geom g, Params prm, const uint8_t* __restrict__ flags, const double* __restrict__ Ucv,
         const double* __restrict__ Unew, const int* __restrict__ any_flag,
         double* __restrict__ theta1) {
  SPD_SKIP(g);
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.Ngc) return;
  int gx[3] = {0, 0, 0};
  grid_coords<NDIM>(tid, g.Ng, g.Ngc, gx);
  int c[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < NDIM; ++a) c[a] = gx[a] - NGH;
  if (prm.blending) {
    // a cell whose 3^NDIM neighbourhood is all active cells reads the flags
    // directly: every neighbour and rule applies (the same values and maximum
    // as below)
    bool inner = true;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) inner = inner && c[a] >= 1 && c[a] <= g.Ns[a] - 2;
    if (inner) {
      const long long i0 = cell_index<NDIM>(g, c);
      long long st[3] = {1, 1, 1};
#pragma unroll
      for (int a = 1; a < NDIM; ++a) st[a] = st[a - 1] * g.Ns[a - 1];
      auto F = [&](const int* o) -> double {
        long long d = 0;
#pragma unroll
        for (int a = 0; a < NDIM; ++a) d += o[a] * st[a];
        return (double)flags[i0 + d];
      };
      const int o0[3] = {0, 0, 0};
      double th = F(o0);
#pragma unroll
      for (int a = 0; a < NDIM; ++a)
#pragma unroll
        for (int s = -1; s <= 1; s += 2) {
          int o[3] = {0, 0, 0};
          o[a] = s;
          th = nmax(0.75 * F(o), th);
        }
      if constexpr (NDIM >= 2) {
        const BlendRule* rules = (NDIM == 2) ? kRules2 : kRules3;
        const int nr = (NDIM == 2) ? 4 : 16;
        for (int k = 0; k < nr; ++k) {
          const BlendRule R = rules[k];
          const int o[3] = {R.off[0], R.off[1], R.off[2]};
          th = nmax(R.w * F(o), th);
        }
      }
      theta1[tid] = th;
      return;
    }
  }
  const bool gate = (*any_flag != 0) && prm.p > 1 && prm.sed;
  const double* Ulast = gate ? Unew : Ucv;  // density is variable 0 of both sets
  double th = trouble_at<NDIM>(g, flags, Ulast, c);
  if (prm.blending) {
    // first neighbours
#pragma unroll
    for (int a = 0; a < NDIM; ++a)
#pragma unroll
      for (int s = -1; s <= 1; s += 2) {
        const int x = gx[a] + s;
        if (x < 0 || x >= g.Ng[a]) continue;
        int cc[3] = {c[0], c[1], c[2]};
        cc[a] += s;
        th = nmax(0.75 * trouble_at<NDIM>(g, flags, Ulast, cc), th);
      }
    if constexpr (NDIM >= 2) {
      const BlendRule* rules = (NDIM == 2) ? kRules2 : kRules3;
      const int nr = (NDIM == 2) ? 4 : 16;
      for (int k = 0; k < nr; ++k) {
        const BlendRule R = rules[k];
        bool ok = true;
#pragma unroll
        for (int a = 0; a < NDIM; ++a) ok = ok && in_slice(gx[a], g.Ng[a], R.tgt[a]);
        if (!ok) continue;
        int cc[3] = {c[0], c[1], c[2]};
#pragma unroll
        for (int a = 0; a < NDIM; ++a) cc[a] += R.off[a];
        th = nmax(R.w * trouble_at<NDIM>(g, flags, Ulast, cc), th);
      }
    }
  }
  theta1[tid] = th;
}

// One pass of the final 0.25 ring along a line of the ghosted grid
// (src/trouble_detection.py:166-169): two sequential vectorised updates.
__device__ __forceinline__ double ring1(int i, int n, double lo, double mid, double hi) {
  const double base = (i <= n - 2) ? nmax(0.25 * (hi > 0.0 ? 1.0 : 0.0), mid) : mid;
  if (i < 1) return base;
  const double prev = nmax(0.25 * (mid > 0.0 ? 1.0 : 0.0), lo);
  return nmax(0.25 * (prev > 0.0 ? 1.0 : 0.0), base);
}

template <int NDIM>
__global__ void __launch_bounds__(256)
k_ring(Geom g, const double* __restrict__ t1, double* __restrict__ theta) {
  SPD_SKIP(g);
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.Ngc) return;
  int gx[3] = {0, 0, 0};
  grid_coords<NDIM>(tid, g.Ng, g.Ngc, gx);
  const int nx = g.Ng[0], ny = g.Ng[1], nz = g.Ng[2];
  auto at = [&](int i, int j, int k) -> double {
    return t1[((long long)k * ny + j) * nx + i];
  };
  auto rx = [&](int j, int k) -> double {
    const int i = gx[0];
    const double lo = (i >= 1) ? at(i - 1, j, k) : 0.0;
    const double hi = (i <= nx - 2) ? at(i + 1, j, k) : 0.0;
    return ring1(i, nx, lo, at(i, j, k), hi);
  };
  double out;
  if constexpr (NDIM == 1) {
    out = rx(0, 0);
  } else if constexpr (NDIM == 2) {
    const int j = gx[1];
    const double lo = (j >= 1) ? rx(j - 1, 0) : 0.0;
    const double hi = (j <= ny - 2) ? rx(j + 1, 0) : 0.0;
    out = ring1(j, ny, lo, rx(j, 0), hi);
  } else {
    auto ry = [&](int k) -> double {
      const int j = gx[1];
      const double lo = (j >= 1) ? rx(j - 1, k) : 0.0;
      const double hi = (j <= ny - 2) ? rx(j + 1, k) : 0.0;
      return ring1(j, ny, lo, rx(j, k), hi);
    };
    const int k = gx[2];
    const double lo = (k >= 1) ? ry(k - 1) : 0.0;
    const double hi = (k <= nz - 2) ? ry(k + 1) : 0.0;
    out = ring1(k, nz, lo, ry(k), hi);
  }
  theta[tid] = out;
}

}  // namespace spd

#include "spd_fv2d.hpp"
#include "spd_pred2d.hpp"

namespace spd {

// ---------------------------------------------------------------------------
// MUSCL-Hancock fallback flux on the FV faces of dimension a whose blending
// coefficient max(theta_L, theta_R) is non-zero, blended into the HO flux in
// place (src/muscl.py:12-35,61-75,129-221, src/fv_simulator.py:88-175,
// src/sdader_simulator.py:325-331).
// ---------------------------------------------------------------------------
// Hancock-predicted primitive state and half slopes of every cell of the
// one-layer-ghosted grid that borders a face with a non-zero blending
// coefficient (or every cell with godunov).  Output layout per cell:
// [Mp (V), S_0 (V), ..., S_{NDIM-1} (V)] on the grid (Ns+2)^NDIM.
template <int NDIM, typename B>
__global__ void __launch_bounds__(256)
k_hancock(Geom g, Params prm, const double* __restrict__ Wold, const double* __restrict__ theta,
          const double* __restrict__ dt_dev, double w_i, B* __restrict__ HS) {
  SPD_SKIP(g);
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
  // needed when the cell is L or R of a face with theta_face > 0 (or NaN)
  bool need = prm.godunov != 0;
  if (!need) {
    long long gi = 0;
#pragma unroll
    for (int b = NDIM - 1; b >= 0; --b) gi = gi * g.Ng[b] + (c[b] + NGH);
    const double th0 = theta[gi];
    need = !(th0 == 0.0);
    long long st = 1;
#pragma unroll
    for (int a = 0; a < NDIM; ++a) {
      if (need) break;
      // faces of this cell along a that exist: the cell must be within the
      // face range of the active transverse cells
      bool trans_ok = true;
#pragma unroll
      for (int b = 0; b < NDIM; ++b)
        if (b != a && (c[b] < 0 || c[b] >= g.Ns[b])) trans_ok = false;
      if (trans_ok) {
        if (c[a] + 1 < g.Ng[a] - NGH && !(theta[gi + st] == 0.0)) need = true;  // face to the right
        if (c[a] - 1 >= -1 && !(theta[gi - st] == 0.0)) need = true;            // face to the left
      }
      st *= g.Ng[a];
    }
  }
  if (!need) return;
  const B dti = (B)((*dt_dev) * w_i);
  B Mp[V], S[NDIM][V];
  {
    // primitives from the FP64 state (FP64 pressure recovery)
    auto load = [&](const int* cc, B* Wb) {
      double Wd[V];
      load_prim<NDIM>(Wold, g, cc, Wd);
#pragma unroll
      for (int v = 0; v < V; ++v) Wb[v] = (B)Wd[v];
    };
    B W[V], Wm[NDIM][V], Wp[NDIM][V];
    load(c, W);
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      int cp[3] = {c[0], c[1], c[2]}, cm[3] = {c[0], c[1], c[2]};
      cp[b] += 1;
      cm[b] -= 1;
      load(cp, Wp[b]);
      load(cm, Wm[b]);
    }
    hancock_state<NDIM, B>(g, prm, c, W, Wm, Wp, dti, Mp, S);
  }
  constexpr int REC = (1 + NDIM) * V;
  B* out = HS + tid;
#pragma unroll
  for (int v = 0; v < V; ++v) out[(long long)v * n1] = Mp[v];
#pragma unroll
  for (int b = 0; b < NDIM; ++b)
#pragma unroll
    for (int v = 0; v < V; ++v) out[(long long)((1 + b) * V + v) * n1] = S[b][v];
  (void)REC;
}

// Fallback flux from the stored Hancock states at the faces of dimension a
// with a non-zero coefficient, blended into the HO flux in place.
template <int NDIM, int RS, int A, typename B>
__global__ void __launch_bounds__(256)
k_fbface(Geom g, Params prm, const B* __restrict__ HS, const double* __restrict__ theta,
         B* __restrict__ Fho) {
  SPD_SKIP(g);
  constexpr int a = A;
  constexpr int V = Dims<NDIM>::V;
  long long ext[3] = {1, 1, 1};
  long long nface = 1;
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
  long long gl = 0, gr = 0, hl = 0, hr = 0, n1 = 1;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) {
    gl = gl * g.Ng[b] + c[b] + NGH - (b == a ? 1 : 0);
    gr = gr * g.Ng[b] + c[b] + NGH;
    hl = hl * (g.Ns[b] + 2) + c[b] + 1 - (b == a ? 1 : 0);
    hr = hr * (g.Ns[b] + 2) + c[b] + 1;
  }
#pragma unroll
  for (int b = 0; b < NDIM; ++b) n1 *= g.Ns[b] + 2;
  double th = nmax(theta[gl], theta[gr]);
  if (prm.godunov) th = 1.0;
  if (th == 0.0) return;
  B WL[V], WR[V], UL[V], UR[V], F[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    WL[v] = HS[(long long)v * n1 + hl] + HS[(long long)((1 + a) * V + v) * n1 + hl];
    WR[v] = HS[(long long)v * n1 + hr] - HS[(long long)((1 + a) * V + v) * n1 + hr];
  }
  const B gam = (B)prm.gamma;
  cons<NDIM>(WL, UL, a, gam);
  cons<NDIM>(WR, UR, a, gam);
  riemann<NDIM, RS>(WL, WR, UL, UR, F, a, gam, (B)prm.min_c2, prm.lowmach_fv);
  const B tb = (B)th;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    B* dst = Fho + v * nface + tid;
    *dst = tb * F[v] + (B(1) - tb) * (*dst);
  }
}

// Maximum wave speed over a primitive subcell-mean array in element layout
// (per-block partial maxima).
template <int NDIM>
__global__ void __launch_bounds__(256)
k_wave_max(Geom g, Params prm, int NL, const double* __restrict__ W, double* __restrict__ partial) {
  SPD_SKIP(g);
  constexpr int P = Dims<NDIM>::P;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npts = g.NE * NL;
  double cw = 0.0;
  if (tid < npts) {
    cw = sound(W[P * npts + tid], W[tid], prm.gamma, prm.min_c2) * NDIM;
#pragma unroll
    for (int k = 1; k <= NDIM; ++k) cw = cw + fabs(W[k * npts + tid]);
  }
  cw = block_max(cw);
  if (threadIdx.x == 0) partial[blockIdx.x] = cw;
}

// Reduce the block partial maxima into *cmax (one block of 1024 threads).
__device__ __forceinline__ void k_reduce_max_body(const BlockIx bi, const double* __restrict__ partial, long long n, double* cmax, const int* skip) {
  if (*skip) return;
  // eight independent loads in flight per thread (the maximum does not depend
  // on the order in which the partials are combined)
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const long long st = blockDim.x;
  long long i = threadIdx.x;
  for (; i + 7 * st < n; i += 8 * st) {
    double q[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) q[j] = partial[i + j * st];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = nmax(a[j], q[j]);
  }
  for (; i < n; i += st) a[0] = nmax(a[0], partial[i]);
  double v = nmax(nmax(nmax(a[0], a[1]), nmax(a[2], a[3])), nmax(nmax(a[4], a[5]), nmax(a[6], a[7])));
  v = block_max(v);
  if (threadIdx.x == 0) *cmax = v;
}
__global__ void __launch_bounds__(1024)
k_reduce_max(const double* __restrict__ partial, long long n, double* cmax, const int* skip) {
  k_reduce_max_body(hw_block(), partial, n, cmax, skip);
}

__global__ void k_swap_scalar(double* a, double* b) {
  if (threadIdx.x == 0) {
    const double t = *a;
    *a = *b;
    *b = t;
  }
}

__global__ void k_dt_final(const double* cmax, double h, double cfl, double* dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dt[0] = cfl * (h / *cmax);
}


__global__ void k_run_init(RunState* st, int* skip, double time, double t_end, long long n_step,
                           long long n_left, int clamp) {
  if (threadIdx.x || blockIdx.x) return;
  st->time = time;
  st->t_end = t_end;
  st->dt = 0.0;
  st->n_step = n_step;
  st->n_left = n_left;
  st->done = (n_left == 0) ? 1 : 0;
  st->clamp = clamp;
  *skip = st->done;
}

// dt of the next step from the recorded maximum wave speed, with the
// reference's clamp to t_end and its dt < 1e-14 stop.
__device__ __forceinline__ void k_run_dt_body(const BlockIx bi, const double* cmax, double h, double cfl, RunState* st, double* dt_out,
                         int* skip) {
  if (threadIdx.x || bi.x) return;
  if (st->done) {
    *skip = 1;
    return;
  }
  if (st->clamp && !(st->time < st->t_end)) {
    st->done = 1;
    *skip = 1;
    return;
  }
  double dt = cfl * (h / *cmax);
  if (st->clamp) {
    if (st->time + dt >= st->t_end) dt = st->t_end - st->time;
    if (dt < 1e-14) {
      st->dt = dt;
      st->done = 1;
      *skip = 1;
      return;
    }
  }
  st->dt = dt;
  *dt_out = dt;
  *skip = 0;
}
__global__ void 
k_run_dt(const double* cmax, double h, double cfl, RunState* st, double* dt_out,
                         int* skip) {
  k_run_dt_body(hw_block(), cmax, h, cfl, st, dt_out, skip);
}

__device__ __forceinline__ void k_run_advance_body(const BlockIx bi, RunState* st, const int* skip) {
  if (threadIdx.x || bi.x) return;
  if (*skip) return;
  st->time = st->time + st->dt;
  st->n_step += 1;
  if (st->n_left > 0) {
    st->n_left -= 1;
    if (st->n_left == 0) st->done = 1;
  }
}
__global__ void 
k_run_advance(RunState* st, const int* skip) {
  k_run_advance_body(hw_block(), st, skip);
}

// Quadrature points of the initial-condition subcell means
// (src/polynomials.py:124-148, src/sd_simulator.py:59-90): for every element,
// local subcell and 2-point Gauss-Legendre node per dimension,
//   pts[a] = low_a + width_a * xq   (separately rounded, as in the reference)
// Layout (ndim, E_{d-1}..E_0, l_{d-1}..l_0, q_{d-1}..q_0); lows / widths hold
// N_a * n values per dimension, concatenated.
template <int NDIM>
__global__ void __launch_bounds__(256)
k_ic_points(Geom g, int n, const double* __restrict__ lows, const double* __restrict__ wids,
            double xq0, double xq1, double* __restrict__ pts) {
  constexpr int Q = 1 << NDIM;
  const long long npts = g.Nc * Q;  // elements * n^d * 2^d
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= npts) return;
  long long x = tid;
  int q[3] = {0, 0, 0}, l[3] = {0, 0, 0}, e[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    q[a] = split_index(x, 2);
  }
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    l[a] = split_index(x, n);
  }
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    e[a] = split_index(x, g.N[a]);
  }
  long long off = 0;
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    const long long k = off + (long long)e[a] * n + l[a];
    const double xq = q[a] ? xq1 : xq0;
    pts[a * npts + tid] = __dadd_rn(lows[k], __dmul_rn(wids[k], xq));
    off += (long long)g.N[a] * n;
  }
}

// W[e, l] = sum over the 2^d nodes of vals[e, l, q] * prod_a w[q_a].
template <int NDIM>
__global__ void __launch_bounds__(256)
k_ic_mean(Geom g, const double* __restrict__ vals, double w0, double w1,
          double* __restrict__ out) {
  constexpr int Q = 1 << NDIM;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.Nc) return;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    double w = 1.0;
#pragma unroll
    for (int a = NDIM - 1; a >= 0; --a) w = w * (((k >> a) & 1) ? w1 : w0);
    acc += vals[tid * Q + k] * w;
  }
  out[tid] = acc;
}

__device__ __forceinline__ void k_zero_counters_body(const BlockIx bi, int* any_flag) {
  if (threadIdx.x == 0 && bi.x == 0) *any_flag = 0;
}
__global__ void 
k_zero_counters(int* any_flag) {
  k_zero_counters_body(hw_block(), any_flag);
}

template <int NDIM>
__global__ void k_prims_points(Params prm, const double* __restrict__ U, double* __restrict__ W,
                               long long count) {
  constexpr int V = Dims<NDIM>::V;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double u[V], w[V];
#pragma unroll
  for (int v = 0; v < V; ++v) u[v] = U[v * count + i];
  prims<NDIM>(u, w, 0, prm.gamma);
#pragma unroll
  for (int v = 0; v < V; ++v) W[v * count + i] = w[v];
}

template <int NDIM>
__global__ void k_cons_points(Params prm, const double* __restrict__ W, double* __restrict__ U,
                              long long count) {
  constexpr int V = Dims<NDIM>::V;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double w[V], u[V];
#pragma unroll
  for (int v = 0; v < V; ++v) w[v] = W[v * count + i];
  cons<NDIM>(w, u, 0, prm.gamma);
#pragma unroll
  for (int v = 0; v < V; ++v) U[v * count + i] = u[v];
}

// Debug tap: flags -> doubles, theta copy, face coefficients.
template <int NDIM>
__global__ void k_tap(Geom g, const uint8_t* __restrict__ flags, const double* __restrict__ theta,
                      double* tr_out, double* th_out, double* aff0, double* aff1, double* aff2) {
  SPD_SKIP(g);
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tr_out && tid < g.Nc) tr_out[tid] = flags[tid];
  if (th_out && tid < g.Ngc) th_out[tid] = theta[tid];
  double* aff[3] = {aff0, aff1, aff2};
#pragma unroll
  for (int a = 0; a < NDIM; ++a) {
    if (!aff[a]) continue;
    long long ext[3] = {1, 1, 1}, nface = 1;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      ext[b] = g.Ns[b] + (b == a ? 1 : 0);
      nface *= ext[b];
    }
    if (tid >= nface) continue;
    int c[3] = {0, 0, 0};
    long long x = tid;
#pragma unroll
    for (int b = 0; b < NDIM; ++b) {
      c[b] = split_index(x, ext[b]);
    }
    long long gl = 0, gr = 0;
#pragma unroll
    for (int b = NDIM - 1; b >= 0; --b) {
      gl = gl * g.Ng[b] + c[b] + NGH - (b == a ? 1 : 0);
      gr = gr * g.Ng[b] + c[b] + NGH;
    }
    aff[a][tid] = nmax(theta[gl], theta[gr]);
  }
}

// Element traces of U_sp at the first (side 0) or last (side 1) flux point
// of dimension a, at every solution point across that side
// (init_sd_Boundaries, src/sdader_simulator.py:415-423): the stored boundary
// values of "ic" / "pressure" sides, layout (V, points) in bc_point's order.
template <int NDIM, int P>
__global__ void __launch_bounds__(256)
k_bc_traces(Ops<P> op, Geom g, int a, int side, const double* __restrict__ U,
            double* __restrict__ out) {
  constexpr int V = Dims<NDIM>::V, n = P + 1, nf = P + 2;
  constexpr int NL = IPow<n, NDIM>::v;
  const long long Pa = g.Nc / g.Ns[a];
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= V * Pa) return;
  const int v = (int)(tid / Pa);
  const long long tp = tid % Pa;
  int ce[3] = {0, 0, 0}, cl[3] = {0, 0, 0};
  long long x = tp;
#pragma unroll
  for (int b = 0; b < NDIM; ++b) {  // the inverse of bc_point: lower dimensions fastest
    if (b == a) continue;
    const int q = split_index(x, g.Ns[b]);
    ce[b] = q / n;
    cl[b] = q % n;
  }
  ce[a] = (side == 0) ? 0 : g.N[a] - 1;
  const long long e = elem_index<NDIM>(g, ce);
  int lb = 0;
#pragma unroll
  for (int b = NDIM - 1; b >= 0; --b) lb = lb * n + ((b == a) ? 0 : cl[b]);
  const int sa = stride_of<n>(a);
  const double* u = U + ((long long)v * g.NE + e) * NL + lb;
  const int fp = (side == 0) ? 0 : nf - 1;
  double acc = 0.0;
#pragma unroll
  for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * u[s * sa];
  out[tid] = acc;
}

}  // namespace spd

// ===========================================================================
// Host side
// ===========================================================================
using namespace spd;

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

static const char* const kKernelClassNames[KC_COUNT] = {
    "k_face_sd", "k_pred", "k_hoflux", "k_tensor", "k_apply", "k_detect", "k_theta1", "k_ring",
    "k_fbface", "k_cv2sp", "k_hancock", "other"};

namespace {


// Typed launcher of one (ndim, p, bulk type) instantiation.
template <int NDIM, int P, typename B> struct Stepper {
  using OpsB = Ops<P, B>;
  static B* Uad(spd_ctx* c) { return (B*)c->Uad_cur; }
  static B* Ff(spd_ctx* c, int a) { return (B*)c->Fface[a]; }
  static B* Fh(spd_ctx* c, int a) { return (B*)c->Fho[a]; }
  static constexpr int V = NDIM + 2, n = P + 1, nf = P + 2, T = P + 1;
  static constexpr int NL = IPow<n, NDIM>::v, NLP = IPow<n, NDIM - 1>::v;

  static_assert(T * NLP <= 256, "one element's lines must fit one block");
  // Elements per block of the predictor update: the largest that fits 256
  // threads and 48 KiB of LDS, or the tuned value (clamped to that maximum).
  static int epb_max() {
    const int lds_per = T * (V * NL + 1) * (int)sizeof(B);
    return std::max(1, std::min(256 / (T * NLP), 49152 / lds_per));
  }
  static int epb(const spd_ctx* c) {
    const int mx = epb_max();
    return c->launch[SPD_LAUNCH_PRED_EPB] > 0 ? std::min(mx, c->launch[SPD_LAUNCH_PRED_EPB]) : mx;
  }

  template <int RS, int A, typename S>
  static void face_generic(spd_ctx* c, const OpsB& op, const S* src, int Tsrc, hipStream_t s) {
    const long long Na = c->g.N[A];
    const long long tot = (long long)Tsrc * (Na + 1) * (c->g.NE / Na) * NLP;
    c->timed(KC_FACE_SD, s, [&] {
      hipLaunchKernelGGL((k_face_sd<NDIM, P, RS, A, S, B>), dim3(grid_for(tot)), dim3(256), 0, s,
                         op, c->g, c->prm, src, Tsrc, Ff(c, A));
    });
  }

  template <int RS, typename S>
  static void faces_rs(spd_ctx* c, const OpsB& op, const S* src, int Tsrc, hipStream_t s) {
    if constexpr (NDIM == 2) {
      using FT = FaceTile2<P>;
      const long long tiles = (long long)((c->g.N[0] + FT::EX - 1) / FT::EX) *
                              ((c->g.N[1] + FT::EY - 1) / FT::EY);
      c->timed(KC_FACE_SD, s, [&] {
        hipLaunchKernelGGL((k_face_sd2d<P, RS, S, B>), dim3((unsigned)(Tsrc * tiles)), dim3(256), 0,
                           s, op, c->g, c->prm, src, Ff(c, 0), Ff(c, 1));
      });
      return;
    }
    if constexpr (NDIM >= 2) {
      constexpr int FPB = 256 / NLP;
      const long long Na = c->g.N[0];
      const long long nblk = (long long)Tsrc * (c->g.NE / Na) * ((Na + 1 + FPB - 1) / FPB);
      c->timed(KC_FACE_SD, s, [&] {
        hipLaunchKernelGGL((k_face_sd_x<NDIM, P, RS, S, B>), dim3((unsigned)nblk), dim3(256), 0,
                           s, op, c->g, c->prm, src, Tsrc, Ff(c, 0));
      });
    }
    if constexpr (NDIM == 1) face_generic<RS, 0>(c, op, src, Tsrc, s);
    if constexpr (NDIM >= 2) face_generic<RS, 1>(c, op, src, Tsrc, s);
    if constexpr (NDIM >= 3) face_generic<RS, 2>(c, op, src, Tsrc, s);
  }
  template <typename S>
  static void faces(spd_ctx* c, const OpsB& op, const S* src, int Tsrc, hipStream_t s) {
    if (c->cfg.riemann_sd != SPD_RS_LLF)
      faces_rs<RS_HLLC>(c, op, src, Tsrc, s);
    else
      faces_rs<RS_LLF>(c, op, src, Tsrc, s);
  }
  // fused 2-D predictor iteration (element faces solved in the kernel).
  // cmp: 1 the first iteration writes its compact form (the increment of
  // one node), 2 the iteration reads the first iterate in that form.
  template <int RS, bool FIRST, bool SRC_CMP, typename S>
  static void pred2d_launch(spd_ctx* c, const OpsB& op, const S* src, int Tsrc, const double* Usp,
                            const double* dt, B* out, int out_cmp, hipStream_t s) {
    using L = pred2d::Layout<P, FIRST>;
    const int ntx = (c->g.N[0] + L::TX - 1) / L::TX, nty = (c->g.N[1] + L::TY - 1) / L::TY;
    c->timed(KC_PRED, s, [&] {
      hipLaunchKernelGGL((pred2d::k_pred2d<P, RS, S, B, FIRST, SRC_CMP>),
                         dim3((unsigned)(ntx * nty)), dim3(L::NTH), (size_t)L::TOTAL * sizeof(B), s,
                         op, c->g, c->prm, src, Tsrc, Usp, dt, out, out_cmp);
    });
  }
  template <int RS, typename S>
  static void pred2d_rs(spd_ctx* c, const OpsB& op, const S* src, int Tsrc, const double* Usp,
                        const double* dt, B* out, int cmp, hipStream_t s) {
    if (Tsrc == 1)
      pred2d_launch<RS, true, false>(c, op, src, Tsrc, Usp, dt, out, cmp == 1 ? 1 : 0, s);
    else if (cmp == 2)
      pred2d_launch<RS, false, true>(c, op, src, Tsrc, Usp, dt, out, 0, s);
    else
      pred2d_launch<RS, false, false>(c, op, src, Tsrc, Usp, dt, out, 0, s);
  }
  template <typename S>
  static void pred2d(spd_ctx* c, const OpsB& op, const S* src, int Tsrc, const double* Usp,
                     const double* dt, B* out, int cmp, hipStream_t s) {
    if (c->cfg.riemann_sd != SPD_RS_LLF)
      pred2d_rs<RS_HLLC>(c, op, src, Tsrc, Usp, dt, out, cmp, s);
    else
      pred2d_rs<RS_LLF>(c, op, src, Tsrc, Usp, dt, out, cmp, s);
  }

  template <typename S>
  static void pred(spd_ctx* c, const OpsB& op, const S* src, int Tsrc, const double* Usp,
                   const double* dt, B* out_ad, double* out_sp, int mode, hipStream_t s) {
    const int E = epb(c);
    const unsigned blocks = (unsigned)((c->g.NE + E - 1) / E);
    const size_t lds = (size_t)E * T * (V * NL + 1) * sizeof(B);
    c->timed(KC_PRED, s, [&] {
      hipLaunchKernelGGL((k_pred<NDIM, P, S, B>), dim3(blocks), dim3(E * T * NLP), lds, s, op,
                         c->g, c->prm, src, Tsrc, Ff(c, 0), Ff(c, 1), Ff(c, 2), Usp, dt, out_ad,
                         out_sp, mode, E);
    });
  }
  template <int RS>
  static void fb_rs(spd_ctx* c, const double* dt, double w, hipStream_t s) {
    long long n1 = 1;
    for (int b = 0; b < NDIM; ++b) n1 *= c->g.Ns[b] + 2;
    c->timed(KC_HANCOCK, s, [&] {
      hipLaunchKernelGGL((k_hancock<NDIM, B>), dim3(grid_for(n1)), dim3(256), 0, s, c->g, c->prm,
                         c->Wold, c->theta, dt, w, (B*)c->HS);
    });
    fbface_one<RS, 0>(c, s);
    if constexpr (NDIM >= 2) fbface_one<RS, 1>(c, s);
    if constexpr (NDIM >= 3) fbface_one<RS, 2>(c, s);
  }
  template <int RS, int A> static void fbface_one(spd_ctx* c, hipStream_t s) {
    long long nface = 1;
    for (int b = 0; b < NDIM; ++b) nface *= c->g.Ns[b] + (b == A ? 1 : 0);
    c->timed(KC_FBFLUX, s, [&] {
      hipLaunchKernelGGL((k_fbface<NDIM, RS, A, B>), dim3(grid_for(nface)), dim3(256), 0, s, c->g,
                         c->prm, (const B*)c->HS, c->theta, Fh(c, A));
    });
  }

  template <int A>
  static void hoflux_generic(spd_ctx* c, const OpsB& op, int i, hipStream_t s) {
    c->timed(KC_HOFLUX, s, [&] {
      if constexpr (NDIM == 3)
        hipLaunchKernelGGL((k_hoflux3<P, A, B>),
                           dim3(grid_for((c->g.NE * nf + 64 / n - 1) / (64 / n) * 64)), dim3(256), 0,
                           s, op, c->g, c->prm, i, Uad(c), Ff(c, A), Fh(c, A));
      else
        hipLaunchKernelGGL((k_hoflux<NDIM, P, A, B>), dim3(grid_for(c->g.NE * nf)), dim3(256), 0,
                           s, op, c->g, c->prm, i, Uad(c), Ff(c, A), Fh(c, A));
    });
  }
  static void apply(spd_ctx* c, const double* dt, double w, double* out, hipStream_t s) {
    c->timed(KC_APPLY, s, [&] {
      hipLaunchKernelGGL((k_apply<NDIM, B>), dim3(grid_for(c->g.Nc)), dim3(256), 0, s, c->g,
                         c->Ucv, Fh(c, 0), Fh(c, 1), Fh(c, 2), dt, w, out);
    });
  }

  // Form of the 2-D FV substep: 1 k_ho2d + k_apply and the fallback on cell
  // tiles, both over the high-order face-flux arrays; 2 the fused high-order
  // flux + candidate kernel writing the face-flux arrays and the fallback
  // marching column strips over them; 3 the fused kernel without face-flux
  // arrays and the fallback on the compacted list of active tiles, which
  // forms the high-order fluxes of its tiles again; 4 (default) adaptive:
  // per substep the path of 2 or of 3, chosen on the device from the number
  // of tiles the previous substep listed (that of the previous step for a
  // step's first substep; none before the first step).  Forms 2 and 3 give
  // the same bits (one arithmetic for the fallback, no contraction into
  // fused multiply-adds there), so the choice never changes a result.  At
  // Kelvin-Helmholtz 1024^2 p=3, FP64, form 3
  // costs 15.4 against 12.9 ms per step in the first steps (99% of the tiles
  // active: the tile kernel forms every flux of a tile again) and 9.7 against
  // 10.6 ms after 220 steps (19% active: no face-flux arrays written).
  static int fb_form(const spd_ctx* c) {
    return c->launch[SPD_LAUNCH_FB_FORM] > 0 ? c->launch[SPD_LAUNCH_FB_FORM] : 4;
  }
  // Active-tile fraction above which the adaptive form takes the column
  // strips.  Kelvin-Helmholtz 1024^2 p=3 FP64, steps 10-210 from the seeded
  // noise, ms per step: 0.35 11.03, 0.5 10.92, 0.65 10.89, 0.8 10.95,
  // 0.9 10.98.
  static constexpr double kDenseTiles = 0.65;
  static bool fused_ho(const spd_ctx* c) { return !c->cfg.FB || fb_form(c) != 1; }
  // Elements per block of k_ho2d (LDS slab of the block's elements, at most
  // 48 KiB).
  static int ho_epb_max(const spd_ctx*) {
    return std::max(1, 49152 / (V * fv2d::ho_nlp(P) * (int)sizeof(B)));
  }
  static int ho_epb(const spd_ctx* c) {
    const int mx = ho_epb_max(c);
    // default: the element count of a 24 KiB FP64 slab in both forms; p = 3:
    // 64 elements (a sweep at Kelvin-Helmholtz 1024^2 p=3 over 16..80 in both
    // precisions: 64 is the fastest, its blocks cover whole 64-element row
    // segments)
    const int dflt = (P == 3) ? 64 : std::max(1, 24576 / (V * NL * 8));
    return c->launch[SPD_LAUNCH_HO_EPB] > 0 ? std::min(mx, c->launch[SPD_LAUNCH_HO_EPB])
                                            : std::min(mx, dflt);
  }
  // Elements per block of the fused k_hocand2d (at most HocTile<P>::E).
  static int hoc_epb(const spd_ctx* c) {
    constexpr int mx = fv2d::HocTile<P>::E;
    return c->launch[SPD_LAUNCH_HOC_EPB] > 0 ? std::min(mx, c->launch[SPD_LAUNCH_HOC_EPB]) : mx;
  }
  static int launch_value(const spd_ctx* c, int key) {
    return key == SPD_LAUNCH_PRED_EPB ? epb(c)
           : key == SPD_LAUNCH_HO_EPB ? ho_epb(c)
           : key == SPD_LAUNCH_HOC_EPB ? hoc_epb(c)
                                       : fb_form(c);
  }
  static int launch_max(const spd_ctx* c, int key) {
    return key == SPD_LAUNCH_PRED_EPB ? epb_max()
           : key == SPD_LAUNCH_HO_EPB ? ho_epb_max(c)
           : key == SPD_LAUNCH_HOC_EPB ? fv2d::HocTile<P>::E
                                       : 4;
  }

  // One FV substep with the 2-D tiled kernels; the corrected state goes to the
  // other state buffer and the buffers are swapped.
  static void substep2d(spd_ctx* c, const OpsB& op, int i, double w, const double* dt,
                        hipStream_t s) {
    tape_state(c, i, s);
    const int form = fb_form(c);
    const bool fused = fused_ho(c);
    const bool adaptive = c->cfg.FB && form == 4;
    using FT = fv2d::FbTile<P>;
    const long long ntiles = fv2d::fbtile_count<P>(c->g);
    if (fused) {
      // high-order fluxes and candidate in one pass (in place without the
      // fallback), with the first half of the trouble detection
      using HT = fv2d::HocTile<P>;
      double* out = c->cfg.FB ? c->Unew : c->Ucv;
      const bool det = c->cfg.FB && !c->tape_replay;
      const size_t lds = det ? (size_t)HT::template lds_bytes_det<B>(c->need_gate ? 4 : 2)
                             : (size_t)HT::template lds_bytes<B>();
      const int Eb = hoc_epb(c);
      const unsigned hblocks = (unsigned)((long long)c->g.N[1] * ((c->g.N[0] + Eb - 1) / Eb));
      if (det && c->need_gate)
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, c->any_flag);
      if (adaptive)
        hipLaunchKernelGGL(fv2d::k_fb_decide, dim3(1), dim3(64), 0, s, c->g, c->fbt_counters,
                           (long long)(kDenseTiles * (double)ntiles), c->tape_replay ? 1 : 0);
      const int* gate = adaptive ? c->fbt_counters + 2 : nullptr;
      c->timed(KC_HOFLUX, s, [&] {
        const bool wf = c->cfg.FB && (form == 2 || form == 4);
        B* fx = wf ? Fh(c, 0) : nullptr;
        B* fy = wf ? Fh(c, 1) : nullptr;
        // blocks all full: constant index arithmetic
        const bool full = Eb == HT::E && c->g.N[0] % HT::E == 0;
        auto go = [&](auto det_c, auto full_c) {
          constexpr bool D = decltype(det_c)::value, F = decltype(full_c)::value;
          hipLaunchKernelGGL((fv2d::k_hocand2d<P, B, D, F>), dim3(hblocks), dim3(256), lds, s, op,
                             c->g, c->prm, i, Uad(c), Ff(c, 0), Ff(c, 1), c->Ucv, dt, w, out, fx,
                             fy, D ? c->flags : nullptr,
                             (D && c->need_gate) ? c->any_flag : nullptr, gate, Eb);
        };
        using T_ = std::true_type;
        using F_ = std::false_type;
        if (det) {
          if (full) go(T_{}, T_{}); else go(T_{}, F_{});
        } else {
          if (full) go(F_{}, T_{}); else go(F_{}, F_{});
        }
      });
      if (!c->cfg.FB) return;
    } else {
      const int E = ho_epb(c);
      const unsigned blocks = (unsigned)((c->g.NE + E - 1) / E);
      c->timed(KC_HOFLUX, s, [&] {
        hipLaunchKernelGGL((fv2d::k_ho2d<P, B>), dim3(blocks), dim3(256),
                           (size_t)V * E * fv2d::ho_nlp(P) * sizeof(B), s, op, c->g, c->prm, i,
                           Uad(c), Ff(c, 0), Ff(c, 1), Fh(c, 0), Fh(c, 1), E);
      });
      apply(c, dt, w, c->Unew, s);
    }
    const dim3 tiles((c->g.Ns[0] + fv2d::TX - 1) / fv2d::TX, (c->g.Ns[1] + fv2d::TY - 1) / fv2d::TY);
    if (!c->tape_replay) {
      const dim3 dtiles((c->g.Ns[0] + fv2d::DTX - 1) / fv2d::DTX,
                        (c->g.Ns[1] + fv2d::DTY - 1) / fv2d::DTY);
      if (fused) {
        // the smooth-extrema half of the detection, on tiles with candidates
        if (c->prm.p > 1 && c->prm.sed)
          c->timed(KC_DETECT, s, [&] {
            hipLaunchKernelGGL(fv2d::k_sed2d, dtiles, dim3(fv2d::DNT), 0, s, c->g, c->prm,
                               c->Unew, c->flags);
          });
      } else {
        if (c->need_gate)
          hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, c->any_flag);
        c->timed(KC_DETECT, s, [&] {
          hipLaunchKernelGGL(fv2d::k_detect2d, dtiles, dim3(fv2d::DNT), 0, s, c->g, c->prm,
                             c->Ucv, c->Unew, c->flags, c->need_gate ? c->any_flag : nullptr);
        });
      }
      const dim3 gtiles((c->g.Ng[0] + fv2d::HTX - 1) / fv2d::HTX,
                        (c->g.Ng[1] + fv2d::HTY - 1) / fv2d::HTY);
      c->timed(KC_THETA1, s, [&] {
        fv2d::FbMarks mk{nullptr, FT::TXc, FT::TYc, (c->g.N[0] + FT::EX - 1) / FT::EX,
                         (c->g.N[1] + FT::EY - 1) / FT::EY};
        if (adaptive) mk.marks = c->fbt_marks;
        mk.dirty = c->theta_dirty;
        hipLaunchKernelGGL(fv2d::k_theta2d, gtiles, dim3(fv2d::HNT), 0, s, c->g, c->prm,
                           c->flags, c->Ucv, c->Unew, c->any_flag, c->theta, mk);
      });
    }
    tape_theta(c, i, s);
    tap(c, i, s);
    // the corrected state replaces the candidate in place
    // blocks of the compacted fallback take listed tiles from a counter until
    // the list is done
    auto fbtiles = [&]() {
      const unsigned nb = (unsigned)std::min<long long>(ntiles, 256 * 12);
      c->timed(KC_FBFLUX, s, [&] {
        if (c->cfg.riemann_fv != SPD_RS_LLF)
          hipLaunchKernelGGL((fv2d::k_fbtile2d<P, RS_HLLC, B>), dim3(nb), dim3(256), 0, s, op,
                             c->g, c->prm, i, Uad(c), Ff(c, 0), Ff(c, 1), c->Ucv, c->theta,
                             c->fbt_counters, c->fbt_list, dt, w, c->Unew, adaptive);
        else
          hipLaunchKernelGGL((fv2d::k_fbtile2d<P, RS_LLF, B>), dim3(nb), dim3(256), 0, s, op,
                             c->g, c->prm, i, Uad(c), Ff(c, 0), Ff(c, 1), c->Ucv, c->theta,
                             c->fbt_counters, c->fbt_list, dt, w, c->Unew, adaptive);
      });
    };
    if (form == 3) {
      hipLaunchKernelGGL(fv2d::k_zero2, dim3(1), dim3(64), 0, s, c->fbt_counters);
      c->timed(KC_THETA1, s, [&] {
        hipLaunchKernelGGL((fv2d::k_fbtile_list<P>),
                           dim3((unsigned)((ntiles + fv2d::FBL_WAVES - 1) / fv2d::FBL_WAVES)),
                           dim3(64 * fv2d::FBL_WAVES), 0, s, c->g, c->prm, c->theta,
                           c->fbt_counters, c->fbt_list, ntiles);
      });
      fbtiles();
      std::swap(c->Ucv, c->Unew);
      return;
    }
    if (adaptive && !c->tape_replay) {
      // the list of the tiles theta marked (its length decides the next
      // substep's path); the tile kernel runs on it when this substep took
      // the compacted path
      c->timed(KC_THETA1, s, [&] {
        hipLaunchKernelGGL(fv2d::k_fbmark_list, dim3(grid_for(ntiles)), dim3(256), 0, s, c->g,
                           c->prm, c->fbt_marks, c->fbt_counters, c->fbt_list, ntiles);
      });
      fbtiles();
    }
    c->timed(KC_FBFLUX, s, [&] {
      const bool hllc = c->cfg.riemann_fv != SPD_RS_LLF;
      if (form == 1) {
        if (hllc)
          hipLaunchKernelGGL((fv2d::k_fbcorrect2d<RS_HLLC, B>), tiles, dim3(fv2d::NT), 0, s,
                             c->g, c->prm, c->Ucv, c->theta, Fh(c, 0), Fh(c, 1), dt, w, c->Unew);
        else
          hipLaunchKernelGGL((fv2d::k_fbcorrect2d<RS_LLF, B>), tiles, dim3(fv2d::NT), 0, s,
                             c->g, c->prm, c->Ucv, c->theta, Fh(c, 0), Fh(c, 1), dt, w, c->Unew);
      } else {
        const int rows = fv2d::march_rows(c->g.Ns[0], c->g.Ns[1]);
        const dim3 strips((c->g.Ns[0] + fv2d::MOUT - 1) / fv2d::MOUT,
                          (c->g.Ns[1] + rows * fv2d::MWAVES - 1) / (rows * fv2d::MWAVES));
        auto go = [&](auto rs, auto st) {
          hipLaunchKernelGGL((fv2d::k_fbmarch2d<decltype(rs)::value, B, decltype(st)::value>),
                             strips, dim3(fv2d::MW * fv2d::MWAVES), 0, s, c->g, c->prm, c->Ucv,
                             c->theta, Fh(c, 0), Fh(c, 1), dt, w, rows, c->Unew,
                             adaptive ? c->fbt_counters + 2 : nullptr);
        };
        using HL = std::integral_constant<int, RS_HLLC>;
        using LF = std::integral_constant<int, RS_LLF>;
        if (c->g.stored) {
          if (hllc) go(HL{}, std::true_type{}); else go(LF{}, std::true_type{});
        } else {
          if (hllc) go(HL{}, std::false_type{}); else go(LF{}, std::false_type{});
        }
      }
    });
    std::swap(c->Ucv, c->Unew);
  }

  // Tape of the reverse-mode derivative: the FV state entering substep i,
  // and the substep's blending coefficients (recorded, or replayed into
  // c->theta in place of the detection).
  static void tape_state(spd_ctx* c, int i, hipStream_t s) {
    if (!c->tape_on) return;
    const size_t bytes = (size_t)V * c->g.Nc * sizeof(double);
    (void)hipMemcpyAsync(c->tape_ucv + (long long)i * V * c->g.Nc, c->Ucv, bytes,
                         hipMemcpyDeviceToDevice, s);
  }
  static void tape_theta(spd_ctx* c, int i, hipStream_t s) {
    if (!c->tape_on && !c->tape_replay) return;
    const size_t bytes = (size_t)c->g.Ngc * sizeof(double);
    double* slot = c->tape_theta + (long long)i * c->g.Ngc;
    if (c->tape_replay) {
      (void)hipMemcpyAsync(c->theta, slot, bytes, hipMemcpyDeviceToDevice, s);
      if (c->theta_dirty)  // theta no longer matches the tiles' marks
        hipLaunchKernelGGL(fv2d::k_fill_u8, dim3((unsigned)((c->theta_tiles + 255) / 256)),
                           dim3(256), 0, s, c->theta_dirty, c->theta_tiles, (uint8_t)1);
    } else
      (void)hipMemcpyAsync(slot, c->theta, bytes, hipMemcpyDeviceToDevice, s);
  }

  static bool tapping(const spd_ctx* c, int i) {
    return i == c->dbg_sub || c->dbg_sub == SPD_TAP_ALL;
  }
  static void tap(spd_ctx* c, int i, hipStream_t s) {
    if (!tapping(c, i)) return;
    const long long k = (c->dbg_sub == SPD_TAP_ALL) ? i : 0;
    double* aff_o[3] = {nullptr, nullptr, nullptr};
    long long mx = std::max(c->g.Nc, c->g.Ngc);
    for (int a = 0; a < NDIM; ++a) {
      long long nface = 1;
      for (int b = 0; b < NDIM; ++b) nface *= c->g.Ns[b] + (b == a ? 1 : 0);
      if (c->dbg_aff[a]) aff_o[a] = c->dbg_aff[a] + k * nface;
      mx = std::max(mx, nface);
    }
    hipLaunchKernelGGL((k_tap<NDIM>), dim3(grid_for(mx)), dim3(256), 0, s, c->g, c->flags,
                       c->theta, c->dbg_tr ? c->dbg_tr + k * c->g.Nc : nullptr,
                       c->dbg_th ? c->dbg_th + k * c->g.Ngc : nullptr, aff_o[0], aff_o[1],
                       aff_o[2]);
  }

  static int step(spd_ctx* c, double* Usp, double* Ucv_sd, double* Wcv_sd, const double* dt,
                  int snapshot, hipStream_t s) {
    const Ops<P> op = make_ops<P>(*c->ops);
    const OpsB opb = make_ops<P, B>(*c->ops);
    if (!c->g.skip) c->g.skip = c->zero_flag;
    const int m = c->cfg.m;
    // --- ADER predictor (src/sdader_simulator.py:160-200)
    // Output buffer of Picard iteration `it` (0-based): the tape keeps every
    // iterate; otherwise the fused 2-D iterations alternate between two
    // buffers and the generic update works in place.
    const long long it_size = (long long)T * V * c->g.NE * NL;
    auto dst = [&](int it) -> B* {
      if (c->tape_on) return (B*)c->tape_uad + it * it_size;
      if constexpr (NDIM == 2) return (it & 1) ? (B*)c->Uad2 : (B*)c->Uad;
      return (B*)c->Uad;
    };
    if constexpr (NDIM == 2) {
      // the face kernels run once, on the final state, for the corrector
      // the first iterate in compact form when another iteration follows and
      // no tape records the iterates
      const bool cmp = m >= 2 && !c->tape_on;
      pred2d(c, opb, (const double*)Usp, 1, Usp, dt, dst(0), cmp ? 1 : 0, s);
      for (int it = 1; it < m; ++it)
        pred2d(c, opb, (const B*)dst(it - 1), T, Usp, dt, dst(it), (cmp && it == 1) ? 2 : 0, s);
    } else {
      faces(c, opb, (const double*)Usp, 1, s);
      pred(c, opb, (const double*)Usp, 1, Usp, dt, dst(0), nullptr, 0, s);
      for (int it = 1; it < m; ++it) {
        faces(c, opb, (const B*)dst(it - 1), T, s);
        pred(c, opb, (const B*)dst(it - 1), T, Usp, dt, dst(it), nullptr, 0, s);
      }
    }
    c->Uad_cur = dst(m - 1);
    faces(c, opb, (const B*)Uad(c), T, s);
    if (c->cfg.update == SPD_UPDATE_SD) {
      pred(c, opb, (const B*)Uad(c), T, Usp, dt, nullptr, Usp, 1, s);
      return hipGetLastError() == hipSuccess ? SPD_OK : fail(SPD_E_HIP, "kernel launch");
    }
    // --- FV corrector (src/sdader_simulator.py:288-349)
    c->timed(KC_TENSOR, s, [&] {
      if constexpr (NDIM == 2)
        hipLaunchKernelGGL((k_tensor2d<P, 1>), dim3((unsigned)t2d_blocks(c->g, P)), dim3(256), 0,
                           s, op, c->g, c->prm, Usp, c->Ucv, (double*)nullptr, (double*)nullptr,
                           (double*)nullptr);
      else if constexpr (NDIM == 1)
        hipLaunchKernelGGL((k_tensor_el<NDIM, P>), dim3(grid_for(c->g.NE * V)), dim3(256), 0, s,
                           op, c->g, 1, Usp, 0, c->Ucv, 1, (double*)nullptr);
      else
        hipLaunchKernelGGL((k_tensor3d<P, 1>),
                           dim3((unsigned)((c->g.NE + t3d_eb(P) - 1) / t3d_eb(P))), dim3(256), 0, s,
                           op, c->g, c->prm, Usp, c->Ucv, (double*)nullptr, (double*)nullptr,
                           (double*)nullptr);
    });
    for (int i = 0; i < T; ++i) {
      const double w = c->ops->w_tp[i];
      if constexpr (NDIM == 2) {
        substep2d(c, opb, i, w, dt, s);
        continue;
      }
      hoflux_generic<0>(c, opb, i, s);
      if constexpr (NDIM >= 2) hoflux_generic<1>(c, opb, i, s);
      if constexpr (NDIM >= 3) hoflux_generic<2>(c, opb, i, s);
      tape_state(c, i, s);
      if (c->cfg.FB) {
        apply(c, dt, w, c->Unew, s);
        c->timed(KC_DETECT, s, [&] {
          hipLaunchKernelGGL((k_prims_cells<NDIM>), dim3(grid_for(c->g.Nc)), dim3(256), 0, s,
                             c->g, c->prm, c->Ucv, c->Wold);
        });
        if (!c->tape_replay) {
          if (c->need_gate)
            hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, c->any_flag);
          c->timed(KC_DETECT, s, [&] {
            if (c->prm.p > 1 && c->prm.sed)
              hipLaunchKernelGGL((k_rho_p<NDIM>), dim3(grid_for(c->g.Nc)), dim3(256), 0, s,
                                 c->g, c->prm, c->Unew, c->rp);
            hipLaunchKernelGGL((k_detect<NDIM>), dim3(grid_for(c->g.Nc)), dim3(256), 0, s, c->g,
                               c->prm, c->Wold, c->Unew, c->rp, c->flags,
                               c->need_gate ? c->any_flag : nullptr);
          });
          double* th_out = c->prm.blending ? c->theta1 : c->theta;
          c->timed(KC_THETA1, s, [&] {
            hipLaunchKernelGGL((k_theta1<NDIM>), dim3(grid_for(c->g.Ngc)), dim3(256), 0, s, c->g,
                               c->prm, c->flags, c->Ucv, c->Unew, c->any_flag, th_out);
          });
          if (c->prm.blending)
            c->timed(KC_RING, s, [&] {
              hipLaunchKernelGGL((k_ring<NDIM>), dim3(grid_for(c->g.Ngc)), dim3(256), 0, s, c->g,
                                 c->theta1, c->theta);
            });
        }
        tape_theta(c, i, s);
        tap(c, i, s);
        if (c->cfg.riemann_fv != SPD_RS_LLF)
          fb_rs<RS_HLLC>(c, dt, w, s);
        else
          fb_rs<RS_LLF>(c, dt, w, s);
      }
      apply(c, dt, w, c->Ucv, s);
    }
    unsigned nb = grid_for(c->g.NE * NL);
    if constexpr (NDIM == 2) {
      nb = (unsigned)t2d_blocks(c->g, P);
      c->timed(KC_CV2SP, s, [&] {
        hipLaunchKernelGGL((k_tensor2d<P, 0>), dim3(nb), dim3(256), 0, s, op, c->g, c->prm,
                           c->Ucv, Usp, snapshot ? Ucv_sd : nullptr, snapshot ? Wcv_sd : nullptr,
                           c->partial);
      });
    } else if constexpr (NDIM == 1) {
      c->timed(KC_CV2SP, s, [&] {
        hipLaunchKernelGGL((k_tensor_el<NDIM, P>), dim3(grid_for(c->g.NE * V)), dim3(256), 0, s,
                           op, c->g, 0, c->Ucv, 1, Usp, 0, snapshot ? Ucv_sd : nullptr);
      });
      c->timed(KC_CV2SP, s, [&] {
        hipLaunchKernelGGL((k_wave_cells<NDIM>), dim3(nb), dim3(256), 0, s, c->g, c->prm, n,
                           c->Ucv, snapshot ? Wcv_sd : nullptr, c->partial);
      });
    } else {
      nb = (unsigned)((c->g.NE + t3d_eb(P) - 1) / t3d_eb(P));
      c->timed(KC_CV2SP, s, [&] {
        hipLaunchKernelGGL((k_tensor3d<P, 0>), dim3(nb), dim3(256), 0, s, op, c->g, c->prm,
                           c->Ucv, Usp, snapshot ? Ucv_sd : nullptr, snapshot ? Wcv_sd : nullptr,
                           c->partial);
      });
    }
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, c->partial, (long long)nb, c->cmax,
                       c->g.skip);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SPD_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return SPD_OK;
  }

  // the stored element traces of every "ic" / "pressure" side from U_sp
  static int bc_traces(spd_ctx* c, const double* U_sp, hipStream_t s) {
    const Ops<P> op = make_ops<P>(*c->ops);
    for (int a = 0; a < NDIM; ++a)
      for (int sd = 0; sd < 2; ++sd) {
        if (!bc_stored(c->g.bc[a][sd])) continue;
        const long long cnt = (long long)V * (c->g.Nc / c->g.Ns[a]);
        hipLaunchKernelGGL((k_bc_traces<NDIM, P>), dim3(grid_for(cnt)), dim3(256), 0, s, op, c->g,
                           a, sd, U_sp, const_cast<double*>(c->g.bct[a][sd]));
      }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
  }

  static int tensor(spd_ctx* c, int to_cv, const double* in, double* out, hipStream_t s) {
    const Ops<P> op = make_ops<P>(*c->ops);
    if constexpr (NDIM <= 2)
      hipLaunchKernelGGL((k_tensor_el<NDIM, P>), dim3(grid_for(c->g.NE * V)), dim3(256), 0, s, op,
                         c->g, to_cv, in, 0, out, 0, (double*)nullptr);
    else
      hipLaunchKernelGGL((k_tensor<NDIM, P>), dim3(grid_for(c->g.NE * V * NL)), dim3(256), 0, s, op,
                         c->g, to_cv, in, out, 0);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
  }
};

using StepFn = int (*)(spd_ctx*, double*, double*, double*, const double*, int, hipStream_t);
using TensorFn = int (*)(spd_ctx*, int, const double*, double*, hipStream_t);
using TracesFn = int (*)(spd_ctx*, const double*, hipStream_t);
using LaunchFn = int (*)(const spd_ctx*, int);
using LaunchMaxFn = int (*)(const spd_ctx*, int);

struct Entry {
  int ndim, p, precision;
  StepFn step;
  TensorFn tensor;
  LaunchFn launch_value;
  LaunchMaxFn launch_max;
  TracesFn bc_traces;
};

#define SPD_ENTRY(D, P, PR, B)                                              \
  {D, P, PR, &Stepper<D, P, B>::step, &Stepper<D, P, B>::tensor,            \
   &Stepper<D, P, B>::launch_value, &Stepper<D, P, B>::launch_max, &Stepper<D, P, B>::bc_traces}
#define SPD_ENTRIES(D, P) \
  SPD_ENTRY(D, P, SPD_PRECISION_FP64, double), SPD_ENTRY(D, P, SPD_PRECISION_FP32, float)
const Entry kEntries[] = {
    SPD_ENTRIES(1, 1), SPD_ENTRIES(1, 2), SPD_ENTRIES(1, 3), SPD_ENTRIES(1, 4), SPD_ENTRIES(1, 5),
    SPD_ENTRIES(2, 1), SPD_ENTRIES(2, 2), SPD_ENTRIES(2, 3), SPD_ENTRIES(2, 4), SPD_ENTRIES(2, 5),
    SPD_ENTRIES(3, 1), SPD_ENTRIES(3, 2), SPD_ENTRIES(3, 3), SPD_ENTRIES(3, 4), SPD_ENTRIES(3, 5),
};
#undef SPD_ENTRIES
#undef SPD_ENTRY

const Entry* find_entry(int ndim, int p, int precision) {
  for (const Entry& e : kEntries)
    if (e.ndim == ndim && e.p == p && e.precision == precision) return &e;
  return nullptr;
}
size_t bulk_size(int precision) { return precision == SPD_PRECISION_FP32 ? 4 : 8; }

// Scratch of the allocation-free dt (spd_compute_dt_static): a file-scope
// device array, one per GPU, held until the process ends.  Layout: bytes
// [0, 256) the skip flag the kernels test (an int, cleared by every call),
// [256, 512) the maximum wave speed, [512, ...) the per-block maxima of at
// most kStaticPartials blocks of 256 solution points.
constexpr long long kStaticPartials = 1ll << 17;
constexpr size_t kStaticScratchBytes = 512 + (size_t)kStaticPartials * sizeof(double);
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device (the array's name is only a host-side stand-in for it).
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

int static_scratch(unsigned char** out) {
  int dev = 0;
  SPD_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SPD_E_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SPD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SPD_OK;
}

}  // namespace

#include "spd_batch.hpp"

namespace {

int batch_build(spd_batch* b, int par) {
  using namespace spd::batch;
  const bool f32 = b->precision == SPD_PRECISION_FP32;
  switch (b->p) {
    case 1: return f32 ? Plan2D<1, float>::build(b, par) : Plan2D<1, double>::build(b, par);
    case 2: return f32 ? Plan2D<2, float>::build(b, par) : Plan2D<2, double>::build(b, par);
    case 3: return f32 ? Plan2D<3, float>::build(b, par) : Plan2D<3, double>::build(b, par);
    case 4: return f32 ? Plan2D<4, float>::build(b, par) : Plan2D<4, double>::build(b, par);
    case 5: return f32 ? Plan2D<5, float>::build(b, par) : Plan2D<5, double>::build(b, par);
    default: return fail(SPD_E_INVALID, "no batched kernels for this order");
  }
}

}  // namespace

extern "C" {

static int need_boundary_values(const spd_ctx* ctx);

int spd_batch_create(spd_ctx** members, double** U_sp, int count, spd_batch** out) {
  if (!members || !U_sp || !out || count < 1) return fail(SPD_E_INVALID, "invalid argument");
  *out = nullptr;
  const spd_ctx* c0 = members[0];
  for (int i = 0; i < count; ++i) {
    const spd_ctx* c = members[i];
    if (!c || !U_sp[i]) return fail(SPD_E_INVALID, "null member or state");
    if (c->ndim != 2 || c->cfg.update != SPD_UPDATE_FV || !c->cfg.FB)
      return fail(SPD_E_INVALID, "batched steps need 2-D members with update=FV and the fallback");
    if (c->p != c0->p || c->cfg.m != c0->cfg.m || c->cfg.precision != c0->cfg.precision ||
        c->cfg.riemann_sd != c0->cfg.riemann_sd || c->cfg.riemann_fv != c0->cfg.riemann_fv ||
        (c->g.stored != 0) != (c0->g.stored != 0) || c->need_gate != c0->need_gate)
      return fail(SPD_E_INVALID,
                  "batch members differ in order, precision, Riemann solvers or side kinds");
    if (c->tape_on || c->tape_replay || c->prof_on || c->dbg_sub != -1)
      return fail(SPD_E_INVALID, "a batch member is taping, profiling or tapping");
    if (need_boundary_values(c) != SPD_OK) return SPD_E_INVALID;
    for (int j = 0; j < i; ++j)
      if (members[j] == c) return fail(SPD_E_INVALID, "a context appears twice in the batch");
  }
  std::unique_ptr<spd_batch> b(new spd_batch());
  b->p = c0->p;
  b->precision = c0->cfg.precision;
  b->T = c0->T;
  for (int i = 0; i < count; ++i) {
    b->mem.push_back(members[i]);
    b->usp.push_back(U_sp[i]);
    b->ucv0.push_back(members[i]->Ucv);
    b->unew0.push_back(members[i]->Unew);
  }
  int rc = batch_build(b.get(), 0);
  if (rc == SPD_OK && (b->T % 2)) rc = batch_build(b.get(), 1);
  if (rc != SPD_OK) return rc;
  *out = b.release();
  return SPD_OK;
}

int spd_batch_run_steps(spd_batch* b, int nsteps, void* stream) {
  if (!b || nsteps < 0) return fail(SPD_E_INVALID, "invalid argument");
  hipStream_t s = (hipStream_t)stream;
  const size_t nm = b->mem.size();
  int par = (b->mem[0]->Ucv == b->ucv0[0]) ? 0 : 1;
  for (size_t i = 0; i < nm; ++i)
    if ((b->mem[i]->Ucv == b->ucv0[i] ? 0 : 1) != par)
      return fail(SPD_E_INVALID, "batch members were stepped apart since the batch was made");
  for (int k = 0; k < nsteps; ++k) {
    for (auto& f : b->plan[par]) f(s);
    if (b->T % 2) par ^= 1;
  }
  const int m = b->mem[0]->cfg.m;
  for (size_t i = 0; i < nm; ++i) {
    spd_ctx* c = b->mem[i];
    c->Ucv = par ? b->unew0[i] : b->ucv0[i];
    c->Unew = par ? b->ucv0[i] : b->unew0[i];
    c->Uad_cur = ((m - 1) & 1) ? c->Uad2 : c->Uad;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_batch_destroy(spd_batch* b) {
  delete b;
  return SPD_OK;
}

const char* spd_last_error(void) { return g_last_error.c_str(); }

int spd_abi_version(void) { return SPD_ABI_VERSION; }

size_t spd_config_size(void) { return sizeof(spd_config); }

int spd_host_operators(int p, int m, double* out, int len) {
  if (!out || p < 0 || m < 0) return fail(SPD_E_INVALID, "invalid argument");
  try {
    host::OpsHost o(p, m);
    std::vector<double> v;
    auto put = [&](const host::Mat& M) {
      for (auto& r : M) v.insert(v.end(), r.begin(), r.end());
    };
    put(o.sp_to_fp);
    put(o.dfp_to_sp);
    put(o.sp_to_cv);
    put(o.cv_to_sp);
    put(o.invader);
    v.insert(v.end(), o.w_tp.begin(), o.w_tp.end());
    v.insert(v.end(), o.x_tp.begin(), o.x_tp.end());
    if ((int)v.size() > len) return fail(SPD_E_INVALID, "buffer too small");
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return (int)v.size();
  } catch (const std::exception& ex) {
    return fail(SPD_E_INVALID, ex.what());
  }
}

int spd_host_geometry(int N, double h, int p, double* out, int len) {
  if (!out || N < 1 || p < 0) return fail(SPD_E_INVALID, "invalid argument");
  host::GeomHost gh(N, h, host::flux_points(p), NGH);
  std::vector<double> v;
  v.insert(v.end(), gh.faces.begin(), gh.faces.end());
  v.insert(v.end(), gh.centers.begin(), gh.centers.end());
  v.insert(v.end(), gh.hfp.begin(), gh.hfp.end());
  v.insert(v.end(), gh.hcv.begin(), gh.hcv.end());
  if ((int)v.size() > len) return fail(SPD_E_INVALID, "buffer too small");
  std::memcpy(out, v.data(), v.size() * sizeof(double));
  return (int)v.size();
}

int spd_ctx_create(const spd_config* cfg, spd_ctx** out) {
  if (!cfg || !out) return fail(SPD_E_INVALID, "null argument");
  *out = nullptr;
  const int nd = cfg->ndim;
  if (nd < 1 || nd > 3) return fail(SPD_E_INVALID, "ndim must be 1, 2 or 3");
  if (cfg->precision != SPD_PRECISION_FP64 && cfg->precision != SPD_PRECISION_FP32)
    return fail(SPD_E_INVALID, "precision must be SPD_PRECISION_FP64 or SPD_PRECISION_FP32");
  if (!find_entry(nd, cfg->p, cfg->precision))
    return fail(SPD_E_UNSUPPORTED, "no kernels built for ndim=" + std::to_string(nd) +
                                       " p=" + std::to_string(cfg->p));
  if (cfg->m != cfg->p) return fail(SPD_E_UNSUPPORTED, "only m == p is built");
  if (cfg->update != SPD_UPDATE_SD && cfg->update != SPD_UPDATE_FV)
    return fail(SPD_E_INVALID, "update must be SD or FV");
  auto rs_ok = [](int r) { return r == SPD_RS_LLF || r == SPD_RS_HLLC || r == SPD_RS_LHLLC; };
  if (!rs_ok(cfg->riemann_sd) || !rs_ok(cfg->riemann_fv))
    return fail(SPD_E_UNSUPPORTED, "riemann solver must be llf, hllc or lhllc");
  if (cfg->slope_limiter != SPD_LIMITER_MINMOD && cfg->slope_limiter != SPD_LIMITER_MONCEN)
    return fail(SPD_E_UNSUPPORTED, "slope limiter must be minmod or moncen");
  for (int a = 0; a < nd; ++a) {
    if (cfg->N[a] < 1) return fail(SPD_E_INVALID, "N must be >= 1");
    if (cfg->N[a] * (cfg->p + 1) < NGH)
      return fail(SPD_E_INVALID, "fewer subcells than ghost layers");
    for (int s = 0; s < 2; ++s)
      if (cfg->bc[a][s] < SPD_BC_PERIODIC || cfg->bc[a][s] > SPD_BC_PRESSURE)
        return fail(SPD_E_UNSUPPORTED, "boundary kind");
    if ((cfg->bc[a][0] == SPD_BC_PRESSURE && cfg->bc[a][1] == SPD_BC_PERIODIC) ||
        (cfg->bc[a][1] == SPD_BC_PRESSURE && cfg->bc[a][0] == SPD_BC_PERIODIC))
      return fail(SPD_E_UNSUPPORTED,
                  "a pressure side opposite a periodic side (the reference's result depends "
                  "on the order in which it stores the two sides)");
    if (!(cfg->lim[a][1] > cfg->lim[a][0])) return fail(SPD_E_INVALID, "empty domain");
  }
  std::unique_ptr<spd_ctx> c(new spd_ctx());
  c->cfg = *cfg;
  SPD_HIP_CHECK(hipGetDevice(&c->device));
  c->ndim = nd;
  c->p = cfg->p;
  c->n = cfg->p + 1;
  c->nf = cfg->p + 2;
  c->T = cfg->m + 1;
  c->V = nd + 2;
  c->NL = 1;
  for (int a = 0; a < nd; ++a) c->NL *= c->n;
  c->NLP = c->NL / c->n;
  try {
    c->ops.reset(new host::OpsHost(cfg->p, cfg->m));
  } catch (const std::exception& ex) {
    return fail(SPD_E_INVALID, ex.what());
  }
  Geom& g = c->g;
  g.NE = 1;
  g.Nc = 1;
  g.Ngc = 1;
  double hmin = 1e10;
  for (int a = 0; a < 3; ++a) {
    const bool on = a < nd;
    g.N[a] = on ? cfg->N[a] : 1;
    g.Ns[a] = on ? cfg->N[a] * c->n : 1;
    g.Ng[a] = on ? g.Ns[a] + 2 * NGH : 1;
    g.bc[a][0] = on ? cfg->bc[a][0] : BC_PERIODIC;
    g.bc[a][1] = on ? cfg->bc[a][1] : BC_PERIODIC;
    const double len = cfg->lim[a][1] - cfg->lim[a][0];
    g.h[a] = on ? len / cfg->N[a] : 1.0;
    if (on) {
      hmin = std::min(hmin, g.h[a]);
      g.NE *= g.N[a];
      g.Nc *= g.Ns[a];
      g.Ngc *= g.Ng[a];
    }
  }
  c->h_dt = hmin / (cfg->p + 1);
  Params& pr = c->prm;
  pr.gamma = cfg->gamma;
  pr.min_c2 = cfg->min_c2;
  pr.tol = cfg->tolerance;
  pr.min_rho = cfg->min_rho;
  pr.max_rho = cfg->max_rho;
  pr.min_P = cfg->min_P;
  pr.p = cfg->p;
  pr.nad_delta = cfg->NAD == SPD_NAD_DELTA;
  pr.sed = cfg->SED;
  pr.pad = cfg->PAD;
  pr.blending = cfg->blending;
  pr.godunov = cfg->godunov;
  pr.lowmach_sd = cfg->riemann_sd == SPD_RS_LHLLC;
  pr.lowmach_fv = cfg->riemann_fv == SPD_RS_LHLLC;
  pr.moncen = cfg->slope_limiter == SPD_LIMITER_MONCEN;

  // geometry arrays
  std::vector<double> gbuf;
  std::vector<size_t> offs[3];
  for (int a = 0; a < nd; ++a) {
    host::GeomHost gh(g.N[a], g.h[a], c->ops->fp, NGH);
    offs[a] = {gbuf.size()};
    gbuf.insert(gbuf.end(), gh.faces.begin(), gh.faces.end());
    offs[a].push_back(gbuf.size());
    gbuf.insert(gbuf.end(), gh.centers.begin(), gh.centers.end());
    offs[a].push_back(gbuf.size());
    gbuf.insert(gbuf.end(), gh.hfp.begin(), gh.hfp.end());
    offs[a].push_back(gbuf.size());
    gbuf.insert(gbuf.end(), gh.hcv.begin(), gh.hcv.end());
    offs[a].push_back(gbuf.size());
    gbuf.insert(gbuf.end(), gh.rhfp.begin(), gh.rhfp.end());
    offs[a].push_back(gbuf.size());
    gbuf.insert(gbuf.end(), gh.rhcv.begin(), gh.rhcv.end());
    offs[a].push_back(gbuf.size());
    gbuf.insert(gbuf.end(), gh.rdc.begin(), gh.rdc.end());
  }
  int rc;
  if ((rc = c->alloc(&c->geom_buf, gbuf.size())) != SPD_OK) return rc;
  SPD_HIP_CHECK(hipMemcpy(c->geom_buf, gbuf.data(), gbuf.size() * sizeof(double),
                          hipMemcpyHostToDevice));
  for (int a = 0; a < 3; ++a) {
    if (a < nd) {
      g.faces[a] = c->geom_buf + offs[a][0];
      g.centers[a] = c->geom_buf + offs[a][1];
      g.hfp[a] = c->geom_buf + offs[a][2];
      g.hcv[a] = c->geom_buf + offs[a][3];
      g.rhfp[a] = c->geom_buf + offs[a][4];
      g.rhcv[a] = c->geom_buf + offs[a][5];
      g.rdc[a] = c->geom_buf + offs[a][6];
    } else {
      g.faces[a] = g.centers[a] = g.hfp[a] = g.hcv[a] = c->geom_buf;
      g.rhfp[a] = g.rhcv[a] = g.rdc[a] = c->geom_buf;
    }
    g.rh[a] = 1.0 / g.h[a];
  }
  // scratch
  const long long V = c->V, T = c->T, NL = c->NL, NLP = c->NLP;
  const size_t bsz = bulk_size(cfg->precision);
  auto alloc_bulk = [&](void** ptr, long long count) {
    unsigned char* q = nullptr;
    const int r = c->alloc(&q, (size_t)count * bsz);
    *ptr = q;
    return r;
  };
  if ((rc = alloc_bulk(&c->Uad, T * V * g.NE * NL)) != SPD_OK) return rc;
  if (nd == 2 && (rc = alloc_bulk(&c->Uad2, T * V * g.NE * NL)) != SPD_OK) return rc;
  c->Uad_cur = c->Uad;
  for (int a = 0; a < nd; ++a) {
    const long long Na = g.N[a];
    if ((rc = alloc_bulk(&c->Fface[a], T * V * (Na + 1) * (g.NE / Na) * NLP)) != SPD_OK)
      return rc;
  }
  if (cfg->update == SPD_UPDATE_FV) {
    for (int a = 0; a < nd; ++a) {
      long long nface = 1;
      for (int b = 0; b < nd; ++b) nface *= g.Ns[b] + (b == a ? 1 : 0);
      if ((rc = alloc_bulk(&c->Fho[a], V * nface)) != SPD_OK) return rc;
    }
    if ((rc = c->alloc(&c->Ucv, (size_t)(V * g.Nc))) != SPD_OK) return rc;
    if (cfg->FB) {
      if ((rc = c->alloc(&c->Unew, (size_t)(V * g.Nc))) != SPD_OK) return rc;
      if ((rc = c->alloc(&c->flags, (size_t)g.Nc)) != SPD_OK) return rc;
      if ((rc = c->alloc(&c->theta1, (size_t)g.Ngc)) != SPD_OK) return rc;
      if ((rc = c->alloc(&c->theta, (size_t)g.Ngc)) != SPD_OK) return rc;
      long long n1 = 1;
      for (int a = 0; a < nd; ++a) n1 *= g.Ns[a] + 2;
      if (nd == 2) {
        // compacted fallback: tiles of 32/(p+1) x 8/(p+1) elements (at least 1)
        const int ex = std::max(1, 32 / c->n), ey = std::max(1, 8 / c->n);
        const long long ntiles =
            (long long)((g.N[0] + ex - 1) / ex) * ((g.N[1] + ey - 1) / ey);
        if ((rc = c->alloc(&c->fbt_counters, 4)) != SPD_OK) return rc;
        if ((rc = c->alloc(&c->fbt_list, (size_t)ntiles)) != SPD_OK) return rc;
        if ((rc = c->alloc(&c->fbt_marks, (size_t)ntiles)) != SPD_OK) return rc;
        SPD_HIP_CHECK(hipMemset(c->fbt_counters, 0, 4 * sizeof(int)));
        SPD_HIP_CHECK(hipMemset(c->fbt_marks, 0, (size_t)ntiles));
        // theta tiles of k_theta2d, all marked as possibly non-zero
        c->theta_tiles = (long long)((g.Ng[0] + fv2d::HTX - 1) / fv2d::HTX) *
                         ((g.Ng[1] + fv2d::HTY - 1) / fv2d::HTY);
        if ((rc = c->alloc(&c->theta_dirty, (size_t)c->theta_tiles)) != SPD_OK) return rc;
        SPD_HIP_CHECK(hipMemset(c->theta_dirty, 1, (size_t)c->theta_tiles));
      }
      if (nd != 2 && (rc = alloc_bulk(&c->HS, (1 + nd) * V * n1)) != SPD_OK) return rc;
      if (nd != 2 && (rc = c->alloc(&c->rp, (size_t)(2 * g.Nc))) != SPD_OK) return rc;
      if (nd != 2 && (rc = c->alloc(&c->Wold, (size_t)(V * g.Nc))) != SPD_OK) return rc;
    }
  }
  // stored boundary values of "ic" / "pressure" sides
  for (int a = 0; a < nd; ++a)
    for (int sd = 0; sd < 2; ++sd)
      if (bc_stored(cfg->bc[a][sd])) c->g.stored = 1;
  if (c->g.stored) {
    for (int a = 0; a < nd; ++a)
      for (int sd = 0; sd < 2; ++sd) {
        double* q = nullptr;
        if ((rc = c->alloc(&q, (size_t)(V * (g.Nc / g.Ns[a])))) != SPD_OK) return rc;
        c->g.bct[a][sd] = q;
      }
    double* w = nullptr;
    if ((rc = c->alloc(&w, (size_t)(V * g.Ngc))) != SPD_OK) return rc;
    c->g.wgh = w;
  }
  if ((rc = c->alloc(&c->any_flag, 1)) != SPD_OK) return rc;
  if ((rc = c->alloc(&c->zero_flag, 1)) != SPD_OK) return rc;
  if ((rc = c->alloc(&c->run_skip, 1)) != SPD_OK) return rc;
  if ((rc = c->alloc(&c->run, 1)) != SPD_OK) return rc;
  if ((rc = c->alloc(&c->run_dt, 1)) != SPD_OK) return rc;
  SPD_HIP_CHECK(hipMemset(c->zero_flag, 0, sizeof(int)));
  SPD_HIP_CHECK(hipMemset(c->run_skip, 0, sizeof(int)));
  SPD_HIP_CHECK(hipMemset(c->run, 0, sizeof(RunState)));
  SPD_HIP_CHECK(hipMemset(c->run_dt, 0, sizeof(double)));
  c->g.skip = c->zero_flag;
  if ((rc = c->alloc(&c->cmax, 1)) != SPD_OK) return rc;
  c->npartial = (g.NE * c->NL + 255) / 256;
  if (nd == 2) c->npartial = std::max(c->npartial, t2d_blocks(g, c->cfg.p));
  if (nd == 3) c->npartial = std::max(c->npartial, (g.NE + t3d_eb(c->cfg.p) - 1) / t3d_eb(c->cfg.p));
  if ((rc = c->alloc(&c->partial, (size_t)c->npartial)) != SPD_OK) return rc;
  SPD_HIP_CHECK(hipMemset(c->any_flag, 0, sizeof(int)));
  SPD_HIP_CHECK(hipMemset(c->cmax, 0, sizeof(double)));
  for (int a = 0; a < nd; ++a)
    for (int sd = 0; sd < 2; ++sd)
      if (cfg->bc[a][sd] != SPD_BC_PERIODIC) c->need_gate = true;
  for (int a = 0; a < 3; ++a) {
    if (!c->Fface[a]) c->Fface[a] = c->Uad;  // unused dims: any valid pointer
    if (!c->Fho[a]) c->Fho[a] = c->Uad;
  }
  *out = c.release();
  return SPD_OK;
}

int spd_ctx_destroy(spd_ctx* ctx) {
  delete ctx;
  return SPD_OK;
}

size_t spd_ctx_scratch_bytes(const spd_ctx* ctx) { return ctx ? ctx->scratch_bytes : 0; }

int spd_get_operators(const spd_ctx* ctx, double* out, int len) {
  if (!ctx || !out) return fail(SPD_E_INVALID, "null argument");
  return spd_host_operators(ctx->cfg.p, ctx->cfg.m, out, len);
}

int spd_compute_dt(spd_ctx* ctx, const double* W_cv, double* dt_dev, void* stream) {
  if (!ctx || !dt_dev) return fail(SPD_E_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (W_cv) {
    const long long npts = ctx->g.NE * ctx->NL;
    const unsigned nb = grid_for(npts);
    switch (ctx->ndim) {
      case 1:
        hipLaunchKernelGGL((k_wave_max<1>), dim3(nb), dim3(256), 0, s, ctx->g, ctx->prm, ctx->NL,
                           W_cv, ctx->partial);
        break;
      case 2:
        hipLaunchKernelGGL((k_wave_max<2>), dim3(nb), dim3(256), 0, s, ctx->g, ctx->prm, ctx->NL,
                           W_cv, ctx->partial);
        break;
      default:
        hipLaunchKernelGGL((k_wave_max<3>), dim3(nb), dim3(256), 0, s, ctx->g, ctx->prm, ctx->NL,
                           W_cv, ctx->partial);
    }
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, ctx->partial, (long long)nb,
                       ctx->cmax, ctx->zero_flag);
  }
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(64), 0, s, ctx->cmax, ctx->h_dt, ctx->cfg.cfl_coeff,
                     dt_dev);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_compute_dt_static(const spd_config* cfg, const double* W_cv, double* dt_dev,
                          void* stream) {
  if (!cfg || !W_cv || !dt_dev) return fail(SPD_E_INVALID, "null argument");
  const int nd = cfg->ndim;
  if (nd < 1 || nd > 3) return fail(SPD_E_INVALID, "ndim must be 1, 2 or 3");
  if (cfg->p < 0) return fail(SPD_E_INVALID, "p must be >= 0");
  long long NE = 1, NL = 1;
  double hmin = 1e10;
  for (int a = 0; a < nd; ++a) {
    if (cfg->N[a] < 1) return fail(SPD_E_INVALID, "N must be >= 1");
    if (!(cfg->lim[a][1] > cfg->lim[a][0])) return fail(SPD_E_INVALID, "empty domain");
    NE *= cfg->N[a];
    NL *= cfg->p + 1;
    hmin = std::min(hmin, (cfg->lim[a][1] - cfg->lim[a][0]) / cfg->N[a]);
  }
  const long long npts = NE * NL;
  const long long nb = (npts + 255) / 256;
  if (nb > kStaticPartials)
    return fail(SPD_E_INVALID, "W_cv holds " + std::to_string(npts) +
                                   " solution points; the allocation-free dt takes at most " +
                                   std::to_string(kStaticPartials * 256));
  unsigned char* base = nullptr;
  int rc = static_scratch(&base);
  if (rc != SPD_OK) return rc;
  int* skip = reinterpret_cast<int*>(base);
  double* cmax = reinterpret_cast<double*>(base + 256);
  double* partial = reinterpret_cast<double*>(base + 512);
  Geom g{};
  g.NE = NE;
  g.skip = skip;
  Params prm{};
  prm.gamma = cfg->gamma;
  prm.min_c2 = cfg->min_c2;
  hipStream_t s = (hipStream_t)stream;
  // the previous call's contents stay in the array: the skip flag the
  // kernels test must read 0; the block maxima and the maximum are written
  // before they are read
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, skip);
  switch (nd) {
    case 1:
      hipLaunchKernelGGL((k_wave_max<1>), dim3((unsigned)nb), dim3(256), 0, s, g, prm, (int)NL,
                         W_cv, partial);
      break;
    case 2:
      hipLaunchKernelGGL((k_wave_max<2>), dim3((unsigned)nb), dim3(256), 0, s, g, prm, (int)NL,
                         W_cv, partial);
      break;
    default:
      hipLaunchKernelGGL((k_wave_max<3>), dim3((unsigned)nb), dim3(256), 0, s, g, prm, (int)NL,
                         W_cv, partial);
  }
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, partial, nb, cmax, skip);
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(64), 0, s, cmax, hmin / (cfg->p + 1),
                     cfg->cfl_coeff, dt_dev);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_swap_wave_max(spd_ctx* ctx, double* slot, void* stream) {
  if (!ctx || !slot) return fail(SPD_E_INVALID, "null argument");
  hipLaunchKernelGGL(k_swap_scalar, dim3(1), dim3(64), 0, (hipStream_t)stream, ctx->cmax, slot);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_set_boundary_values(spd_ctx* ctx, const double* U_sp, const double* W_ghosted,
                            void* stream) {
  if (!ctx || !U_sp || !W_ghosted) return fail(SPD_E_INVALID, "null argument");
  if (!ctx->g.stored)
    return fail(SPD_E_INVALID, "no side of this context keeps stored values (ic / pressure)");
  hipStream_t s = (hipStream_t)stream;
  const int rc = find_entry(ctx->ndim, ctx->p, ctx->cfg.precision)->bc_traces(ctx, U_sp, s);
  if (rc != SPD_OK) return rc;
  SPD_HIP_CHECK(hipMemcpyAsync(const_cast<double*>(ctx->g.wgh), W_ghosted,
                               (size_t)ctx->V * ctx->g.Ngc * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
  ctx->bc_ready = true;
  return SPD_OK;
}

static int need_boundary_values(const spd_ctx* ctx) {
  if (ctx->g.stored && !ctx->bc_ready)
    return fail(SPD_E_INVALID, "an ic / pressure side needs spd_set_boundary_values first");
  return SPD_OK;
}

int spd_step(spd_ctx* ctx, double* U_sp, double* U_cv, double* W_cv, const double* dt_dev,
             int snapshot, void* stream) {
  if (!ctx || !U_sp || !dt_dev) return fail(SPD_E_INVALID, "null argument");
  if (need_boundary_values(ctx) != SPD_OK) return SPD_E_INVALID;
  if (snapshot && ctx->cfg.update == SPD_UPDATE_FV && (!U_cv || !W_cv))
    return fail(SPD_E_INVALID, "snapshot needs U_cv and W_cv");
  const Entry* e = find_entry(ctx->ndim, ctx->p, ctx->cfg.precision);
  return e->step(ctx, U_sp, U_cv, W_cv, dt_dev, snapshot, (hipStream_t)stream);
}

int spd_set_launch(spd_ctx* ctx, int key, int value) {
  if (!ctx || key < 0 || key >= SPD_LAUNCH_COUNT || value < 0)
    return fail(SPD_E_INVALID, "invalid launch key or value");
  const Entry* en = find_entry(ctx->ndim, ctx->p, ctx->cfg.precision);
  if (value > en->launch_max(ctx, key))
    return fail(SPD_E_INVALID, "launch value above its maximum " +
                                   std::to_string(en->launch_max(ctx, key)));
  ctx->launch[key] = value;
  return SPD_OK;
}

int spd_get_launch(const spd_ctx* ctx, int key, int* value, int* max_value) {
  if (!ctx || key < 0 || key >= SPD_LAUNCH_COUNT) return fail(SPD_E_INVALID, "invalid launch key");
  const Entry* en = find_entry(ctx->ndim, ctx->p, ctx->cfg.precision);
  if (value) *value = en->launch_value(ctx, key);
  if (max_value) *max_value = en->launch_max(ctx, key);
  return SPD_OK;
}

int spd_run_begin(spd_ctx* ctx, double time, double t_end, long long n_step, long long max_steps,
                  int clamp, void* stream) {
  if (!ctx) return fail(SPD_E_INVALID, "null argument");
  hipLaunchKernelGGL(k_run_init, dim3(1), dim3(64), 0, (hipStream_t)stream, ctx->run,
                     ctx->run_skip, time, t_end, n_step, max_steps, clamp);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_run_steps(spd_ctx* ctx, double* U_sp, int nsteps, void* stream) {
  if (!ctx || !U_sp || nsteps < 0) return fail(SPD_E_INVALID, "invalid argument");
  if (need_boundary_values(ctx) != SPD_OK) return SPD_E_INVALID;
  hipStream_t s = (hipStream_t)stream;
  const Entry* en = find_entry(ctx->ndim, ctx->p, ctx->cfg.precision);
  for (int k = 0; k < nsteps; ++k) {
    hipLaunchKernelGGL(k_run_dt, dim3(1), dim3(64), 0, s, ctx->cmax, ctx->h_dt, ctx->cfg.cfl_coeff,
                       ctx->run, ctx->run_dt, ctx->run_skip);
    ctx->g.skip = ctx->run_skip;
    const int rc = en->step(ctx, U_sp, nullptr, nullptr, ctx->run_dt, 0, s);
    ctx->g.skip = ctx->zero_flag;
    if (rc != SPD_OK) return rc;
    hipLaunchKernelGGL(k_run_advance, dim3(1), dim3(64), 0, s, ctx->run, ctx->run_skip);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_run_status(spd_ctx* ctx, double* time, long long* n_step, int* done, double* dt,
                   void* stream) {
  if (!ctx) return fail(SPD_E_INVALID, "null argument");
  RunState st;
  SPD_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  SPD_HIP_CHECK(hipMemcpy(&st, ctx->run, sizeof(RunState), hipMemcpyDeviceToHost));
  if (time) *time = st.time;
  if (n_step) *n_step = st.n_step;
  if (done) *done = st.done;
  if (dt) *dt = st.dt;
  return SPD_OK;
}

int spd_sp_to_regular(spd_ctx* ctx, const double* in, double* out, void* stream) {
  if (!ctx || !in || !out) return fail(SPD_E_INVALID, "null argument");
  return find_entry(ctx->ndim, ctx->p, ctx->cfg.precision)
      ->tensor(ctx, 2, in, out, (hipStream_t)stream);
}

int spd_ic_points(spd_ctx* ctx, const double* lows, const double* wids, double* pts,
                  void* stream) {
  if (!ctx || !lows || !wids || !pts) return fail(SPD_E_INVALID, "null argument");
  std::vector<double> xq, wq;
  host::gauss_legendre(2, xq, wq);
  const long long npts = ctx->g.Nc << ctx->ndim;
  hipStream_t s = (hipStream_t)stream;
  switch (ctx->ndim) {
    case 1:
      hipLaunchKernelGGL(k_ic_points<1>, dim3(grid_for(npts)), dim3(256), 0, s, ctx->g, ctx->n,
                         lows, wids, xq[0], xq[1], pts);
      break;
    case 2:
      hipLaunchKernelGGL(k_ic_points<2>, dim3(grid_for(npts)), dim3(256), 0, s, ctx->g, ctx->n,
                         lows, wids, xq[0], xq[1], pts);
      break;
    default:
      hipLaunchKernelGGL(k_ic_points<3>, dim3(grid_for(npts)), dim3(256), 0, s, ctx->g, ctx->n,
                         lows, wids, xq[0], xq[1], pts);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_ic_mean(spd_ctx* ctx, const double* vals, double* out, void* stream) {
  if (!ctx || !vals || !out) return fail(SPD_E_INVALID, "null argument");
  std::vector<double> xq, wq;
  host::gauss_legendre(2, xq, wq);
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = grid_for(ctx->g.Nc);
  switch (ctx->ndim) {
    case 1:
      hipLaunchKernelGGL(k_ic_mean<1>, dim3(nb), dim3(256), 0, s, ctx->g, vals, wq[0], wq[1], out);
      break;
    case 2:
      hipLaunchKernelGGL(k_ic_mean<2>, dim3(nb), dim3(256), 0, s, ctx->g, vals, wq[0], wq[1], out);
      break;
    default:
      hipLaunchKernelGGL(k_ic_mean<3>, dim3(nb), dim3(256), 0, s, ctx->g, vals, wq[0], wq[1], out);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_sp_to_cv(spd_ctx* ctx, const double* in, double* out, void* stream) {
  if (!ctx || !in || !out) return fail(SPD_E_INVALID, "null argument");
  return find_entry(ctx->ndim, ctx->p, ctx->cfg.precision)->tensor(ctx, 1, in, out, (hipStream_t)stream);
}

int spd_cv_to_sp(spd_ctx* ctx, const double* in, double* out, void* stream) {
  if (!ctx || !in || !out) return fail(SPD_E_INVALID, "null argument");
  return find_entry(ctx->ndim, ctx->p, ctx->cfg.precision)->tensor(ctx, 0, in, out, (hipStream_t)stream);
}

int spd_cons_to_prims(spd_ctx* ctx, const double* U, double* W, long long count, void* stream) {
  if (!ctx || !U || !W || count < 0) return fail(SPD_E_INVALID, "invalid argument");
  if (count == 0) return SPD_OK;
  hipStream_t s = (hipStream_t)stream;
  switch (ctx->ndim) {
    case 1:
      hipLaunchKernelGGL((k_prims_points<1>), dim3(grid_for(count)), dim3(256), 0, s, ctx->prm, U,
                         W, count);
      break;
    case 2:
      hipLaunchKernelGGL((k_prims_points<2>), dim3(grid_for(count)), dim3(256), 0, s, ctx->prm, U,
                         W, count);
      break;
    default:
      hipLaunchKernelGGL((k_prims_points<3>), dim3(grid_for(count)), dim3(256), 0, s, ctx->prm, U,
                         W, count);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_prims_to_cons(spd_ctx* ctx, const double* W, double* U, long long count, void* stream) {
  if (!ctx || !U || !W || count < 0) return fail(SPD_E_INVALID, "invalid argument");
  if (count == 0) return SPD_OK;
  hipStream_t s = (hipStream_t)stream;
  switch (ctx->ndim) {
    case 1:
      hipLaunchKernelGGL((k_cons_points<1>), dim3(grid_for(count)), dim3(256), 0, s, ctx->prm, W,
                         U, count);
      break;
    case 2:
      hipLaunchKernelGGL((k_cons_points<2>), dim3(grid_for(count)), dim3(256), 0, s, ctx->prm, W,
                         U, count);
      break;
    default:
      hipLaunchKernelGGL((k_cons_points<3>), dim3(grid_for(count)), dim3(256), 0, s, ctx->prm, W,
                         U, count);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? SPD_OK : fail(SPD_E_HIP, hipGetErrorString(e));
}

int spd_profile_begin(spd_ctx* ctx, int max_launches) {
  if (!ctx || max_launches < 1) return fail(SPD_E_INVALID, "invalid argument");
  for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
  ctx->prof_ev.assign(2 * (size_t)max_launches, nullptr);
  for (auto& e : ctx->prof_ev) SPD_HIP_CHECK(hipEventCreate(&e));
  ctx->prof_cls.assign((size_t)max_launches, KC_OTHER);
  ctx->prof_used = 0;
  ctx->prof_on = true;
  return SPD_OK;
}

int spd_profile_end(spd_ctx* ctx, double* ms_out, long long* count_out, int n_classes) {
  if (!ctx || !ms_out || !count_out) return fail(SPD_E_INVALID, "null argument");
  ctx->prof_on = false;
  for (int k = 0; k < n_classes; ++k) {
    ms_out[k] = 0.0;
    count_out[k] = 0;
  }
  if (ctx->prof_used) SPD_HIP_CHECK(hipEventSynchronize(ctx->prof_ev[2 * ctx->prof_used - 1]));
  for (size_t i = 0; i < ctx->prof_used; ++i) {
    float ms = 0.f;
    SPD_HIP_CHECK(hipEventElapsedTime(&ms, ctx->prof_ev[2 * i], ctx->prof_ev[2 * i + 1]));
    const int k = ctx->prof_cls[i];
    if (k < n_classes) {
      ms_out[k] += ms;
      count_out[k] += 1;
    }
  }
  const int used = (int)ctx->prof_used;
  for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
  ctx->prof_ev.clear();
  ctx->prof_cls.clear();
  ctx->prof_used = 0;
  return used;
}

int spd_kernel_class_count(void) { return KC_COUNT; }

const char* spd_kernel_class_name(int k) {
  return (k >= 0 && k < KC_COUNT) ? kKernelClassNames[k] : "";
}

int spd_set_debug_tap(spd_ctx* ctx, int substep, double* troubles, double* theta, double* aff_x,
                      double* aff_y, double* aff_z) {
  if (!ctx) return fail(SPD_E_INVALID, "null argument");
  if (substep != -1 && !(ctx->cfg.update == SPD_UPDATE_FV && ctx->cfg.FB))
    return fail(SPD_E_INVALID, "the debug tap needs update=FV and FB");
  if (substep >= ctx->T || substep < SPD_TAP_ALL) return fail(SPD_E_INVALID, "substep out of range");
  ctx->dbg_sub = substep;
  ctx->dbg_tr = troubles;
  ctx->dbg_th = theta;
  ctx->dbg_aff[0] = aff_x;
  ctx->dbg_aff[1] = aff_y;
  ctx->dbg_aff[2] = aff_z;
  return SPD_OK;
}

}  // extern "C"
