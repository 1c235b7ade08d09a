// 2-D ADER predictor iteration with the element-face Riemann problems solved
// in the same kernel (src/sdader_simulator.py:160-200,213-240,148-158,
// src/sd_boundary.py:7-76).  A block owns a TX x TY tile of elements and all
// time nodes; one thread per (element, time node, line) in each dimension
// phase:
//   1. load the line, evaluate its two end traces (sp -> fp 0 and fp nf-1);
//      the upper trace goes to LDS, and a few threads evaluate the traces of
//      the one-element halo across the tile's two sides along the axis;
//   2. every element's lower face flux is solved by the element itself (the
//      upper-side neighbour inside the tile reads it from LDS); the elements
//      whose upper neighbour is outside the tile (or the domain) also solve
//      their upper face;
//   3. flux at the interior flux points, derivative back to the solution
//      points, accumulated in the element's LDS slab (as k_pred);
// then the Picard update writes the next predictor state to another buffer
// (neighbouring blocks still read the current one).  The arithmetic of every
// trace, state conversion, Riemann solve and derivative is that of k_face_sd
// and k_pred.  This file is included into spd_step.hip.
#pragma once

namespace spd {
namespace pred2d {

// Elements per block along x and y: all T time nodes per element (Tile), or
// one (the first iteration, whose source U_sp is the same at every node).
template <int P> struct Tile;
template <> struct Tile<1> { static constexpr int X = 8, Y = 8; };
template <> struct Tile<2> { static constexpr int X = 5, Y = 5; };
template <> struct Tile<3> { static constexpr int X = 4, Y = 4; };
template <> struct Tile<4> { static constexpr int X = 3, Y = 3; };
template <> struct Tile<5> { static constexpr int X = 2, Y = 2; };
template <int P> struct TileFirst;
template <> struct TileFirst<1> { static constexpr int X = 16, Y = 8; };
template <> struct TileFirst<2> { static constexpr int X = 9, Y = 9; };
template <> struct TileFirst<3> { static constexpr int X = 8, Y = 8; };
template <> struct TileFirst<4> { static constexpr int X = 7, Y = 6; };
template <> struct TileFirst<5> { static constexpr int X = 6, Y = 5; };

template <int P, bool FIRST> struct Layout {
  static constexpr int V = 4, n = P + 1, nf = P + 2, T = P + 1, NL = n * n;
  static constexpr int TA = FIRST ? 1 : T;           // time nodes evaluated per element
  static constexpr int TX = FIRST ? TileFirst<P>::X : Tile<P>::X;
  static constexpr int TY = FIRST ? TileFirst<P>::Y : Tile<P>::Y;
  static constexpr int TE = TX * TY;
  static constexpr int NTH = TE * TA * n;            // threads per block
  static constexpr int TM = TX > TY ? TX : TY;       // rows of a halo side
  // dudt slab of one (element, time node), padded so that the 32 lanes of a
  // half-wave (p = 3: two elements x 4 nodes x 4 lines) that write or
  // accumulate one point of their lines fall into different LDS banks (at
  // most two lanes per bank); other degrees keep one value of padding
  static constexpr int SLAB = V * NL + (P == 3 ? 5 : 1);
  static constexpr int LINE = TA * n * V;            // traces of one element / halo row
  static constexpr int TRV = TE * TA * n;            // one variable's plane of the traces
  static constexpr int HV = TM * TA * n;             // one variable's plane of a halo side
  // LDS offsets, in values of the bulk type; the traces are stored variable
  // by variable, so that consecutive lanes touch consecutive values
  static constexpr int O_DUDT = 0;                   // [TE][TA][SLAB]
  static constexpr int O_TR = O_DUDT + TE * TA * SLAB;  // [V][TE][TA][n] upper traces, then lower-face fluxes
  static constexpr int O_HL = O_TR + TE * LINE;      // [V][TM][TA][n] upper traces of the lower halo
  static constexpr int O_HU = O_HL + TM * LINE;      // [V][TM][TA][n] lower traces of the upper halo
  static constexpr int TOTAL = O_HU + TM * LINE;
  static_assert(NTH <= 256, "one block per tile");
  static_assert(2 * TM * TA * n <= NTH, "one halo line per thread");
  static_assert(TOTAL * 8 <= 49152, "LDS of the FP64 form");
};

// FIRST: the first iteration, src = U_sp (Tsrc = 1); otherwise Tsrc = T.
// Compact first iterate: the first iteration's Picard update gives every node
// the same increment, out[tn] = U_sp - dt (sum_p inv[tn][p]) dudt, so with
// out_cmp it writes dudt alone (one node, layout [v][e][l]) and the next
// iteration (SRC_CMP) forms each node's value from U_sp and dudt as it loads
// it, with the same operations in the same order (the same values): one node
// of writes and one of reads instead of T each.
template <int P, int RS, typename S, typename B, bool FIRST, bool SRC_CMP>
__device__ __forceinline__ void k_pred2d_body(const BlockIx bi, Ops<P, B> op, Geom g, Params prm, const S* __restrict__ src, int Tsrc,
         const double* __restrict__ Usp, const double* __restrict__ dt_dev, B* __restrict__ out,
         int out_cmp) {
  static_assert(!(FIRST && SRC_CMP), "the first iteration reads U_sp");
  SPD_SKIP(g);
  using L = Layout<P, FIRST>;
  constexpr int V = 4, n = L::n, nf = L::nf, T = L::TA, NL = L::NL;
  constexpr int TX = L::TX, TY = L::TY, SLAB = L::SLAB;
  B* lds = reinterpret_cast<B*>(spd_dyn_lds);
  B* dudt = lds + L::O_DUDT;
  B* sTR = lds + L::O_TR;
  B* sHL = lds + L::O_HL;
  B* sHU = lds + L::O_HU;
  const int Nx = g.N[0], Ny = g.N[1];
  const long long NE = g.NE;
  const int ntx = (Nx + TX - 1) / TX;
  const int bx = (int)(bi.x % ntx), by = (int)(bi.x / ntx);
  const int tid = threadIdx.x;
  const int el = tid / (T * n), rem = tid % (T * n);
  const int t = rem / n, lp = rem % n;
  const int ex = el % TX, ey = el / TX;
  const int gx = bx * TX + ex, gy = by * TY + ey;
  const bool inside = tid < L::NTH && gx < Nx && gy < Ny;
  const bool active = inside && t < Tsrc;
  const B gam = (B)prm.gamma, mc2 = (B)prm.min_c2;

  // line `l` of element (xe, ye) along dimension a at time node tt
  const double dt_src = SRC_CMP ? *dt_dev : 0.0;
  auto load_line = [&](int xe, int ye, int tt, int a, int l, B (*u)[n]) {
    const long long e = (long long)ye * Nx + xe;
    const int lb = (a == 0) ? l * n : l;
    const int sa = (a == 0) ? 1 : n;
    if constexpr (SRC_CMP) {
      // node tt of the first iterate from U_sp and its increment (the first
      // iteration's Picard update, below)
      B cn[L::T];
#pragma unroll
      for (int p = 0; p < L::T; ++p) cn[p] = op.inv[tt][p];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const long long o = ((long long)v * NE + e) * NL + lb;
#pragma unroll
        for (int s = 0; s < n; ++s) {
          const B d0 = (B)src[o + s * sa];
          B acc = B(0);
#pragma unroll
          for (int p = 0; p < L::T; ++p) acc += cn[p] * d0;
          u[v][s] = (B)(Usp[o + s * sa] - (double)acc * dt_src);
        }
      }
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const S* s0 = src + ((long long)(tt * V + v) * NE + e) * NL + lb;
#pragma unroll
        for (int s = 0; s < n; ++s) u[v][s] = (B)s0[s * sa];
      }
    }
  };
  auto trace = [&](B (*u)[n], int fp, B* q) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
      B acc = B(0);
#pragma unroll
      for (int s = 0; s < n; ++s) acc += op.s2f[fp][s] * u[v][s];
      q[v] = acc;
    }
  };
  auto solve = [&](const B* qa, const B* qb, int a, B* F) {
    B WL[V], WR[V], UL[V], UR[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      UL[v] = qa[v];
      UR[v] = qb[v];
    }
    prims<2>(UL, WL, a, gam);
    prims<2>(UR, WR, a, gam);
    riemann<2, RS>(WL, WR, UL, UR, F, a, gam, mc2, prm.lowmach_sd);
  };

  // Both lines of the thread are loaded up front where the extra registers
  // leave the occupancy as it is (FP64 at p = 1 and 3, whose LDS or registers
  // already hold it at 4 and 3 waves per SIMD), so that the y line's latency
  // hides behind the x phase; elsewhere each line is loaded at its phase.
  constexpr bool PREFETCH = sizeof(B) == 8 && (P == 1 || P == 3);
  B ux[V][n], uy[V][n];
  if (active) {
    load_line(gx, gy, t, 0, lp, ux);
    if (PREFETCH) load_line(gx, gy, t, 1, lp, uy);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int ca = (a == 0) ? gx : gy;        // element coordinate along a
    const int la = (a == 0) ? ex : ey;        // tile coordinate along a
    const int lt = (a == 0) ? ey : ex;        // tile coordinate across a (halo row)
    const int TA = (a == 0) ? TX : TY;
    const int Na = (a == 0) ? Nx : Ny;
    const int nbr = (a == 0) ? 1 : TX;        // element offset of the upper neighbour in the tile
    B(&u)[V][n] = (a == 0) ? ux : uy;
    B qL[V], qR[V];
    if (active) {
      if (!PREFETCH && a == 1) load_line(gx, gy, t, 1, lp, uy);
      trace(u, 0, qL);
      trace(u, nf - 1, qR);
      B* d = sTR + (el * T + t) * n + lp;
#pragma unroll
      for (int v = 0; v < V; ++v) d[v * L::TRV] = qR[v];
    }
    // halo traces: side 0 = the element below the tile's first one along a
    // (its upper trace), side 1 = the element above the tile's last element
    // in the domain (its lower trace); wrapped on periodic sides
    {
      const int rows = (a == 0) ? TY : TX;
      const int per = Tsrc * n;
      if (tid < 2 * rows * per) {
        const int side = tid / (rows * per);
        const int r2 = tid % (rows * per);
        const int k = r2 / per, th = (r2 % per) / n, lh = r2 % n;
        const int across = ((a == 0) ? by * TY : bx * TX) + k;   // coordinate across a
        const int Nacross = (a == 0) ? Ny : Nx;
        int along = (side == 0) ? ((a == 0) ? bx * TX : by * TY) - 1
                                : min(((a == 0) ? bx * TX : by * TY) + TA, Na);
        bool ok = across < Nacross;
        if (along < 0 || along >= Na) {
          if (g.bc[a][side] == BC_PERIODIC)
            along = (along < 0) ? Na - 1 : 0;
          else
            ok = false;
        }
        if (ok) {
          B uh[V][n], q[V];
          if (a == 0)
            load_line(along, across, th, a, lh, uh);
          else
            load_line(across, along, th, a, lh, uh);
          trace(uh, side == 0 ? nf - 1 : 0, q);
          B* d = ((side == 0) ? sHL : sHU) + (k * T + th) * n + lh;
#pragma unroll
          for (int v = 0; v < V; ++v) d[v * L::HV] = q[v];
        }
      }
    }
    __syncthreads();
    // lower face (and, at the tile's upper side, the upper face)
    B F[nf][V];
    if (active) {
      B qa[V];
      // transverse solution point of the thread's line (for stored sides)
      const long long tp = (long long)((a == 0) ? gy : gx) * n + lp;
      if (ca == 0 && bc_stored(g.bc[a][0])) {
        bc_stored_traces<V, B>(g, a, 0, tp, qa, qL);
      } else if (ca == 0 && g.bc[a][0] != BC_PERIODIC) {
#pragma unroll
        for (int v = 0; v < V; ++v) qa[v] = qL[v];
        if (g.bc[a][0] == BC_REFLECTIVE) qa[a + 1] = -qa[a + 1];
      } else {
        const B* s = (la > 0) ? sTR + ((el - nbr) * T + t) * n + lp : sHL + (lt * T + t) * n + lp;
        const int sv = (la > 0) ? L::TRV : L::HV;
#pragma unroll
        for (int v = 0; v < V; ++v) qa[v] = s[v * sv];
      }
      solve(qa, qL, a, F[0]);
      if (ca == Na - 1 || la == TA - 1) {
        B qb[V];
        if (ca == Na - 1 && bc_stored(g.bc[a][1])) {
          bc_stored_traces<V, B>(g, a, 1, tp, qb, qR);
        } else if (ca == Na - 1 && g.bc[a][1] != BC_PERIODIC) {
#pragma unroll
          for (int v = 0; v < V; ++v) qb[v] = qR[v];
          if (g.bc[a][1] == BC_REFLECTIVE) qb[a + 1] = -qb[a + 1];
        } else {
          const B* s = sHU + (lt * T + t) * n + lp;
#pragma unroll
          for (int v = 0; v < V; ++v) qb[v] = s[v * L::HV];
        }
        solve(qR, qb, a, F[nf - 1]);
      }
    }
    __syncthreads();  // every upper trace has been read
    if (active) {
      B* d = sTR + (el * T + t) * n + lp;
#pragma unroll
      for (int v = 0; v < V; ++v) d[v * L::TRV] = F[0][v];
    }
    __syncthreads();
    if (active) {
      if (!(ca == Na - 1 || la == TA - 1)) {
        const B* s = sTR + ((el + nbr) * T + t) * n + lp;
#pragma unroll
        for (int v = 0; v < V; ++v) F[nf - 1][v] = s[v * L::TRV];
      }
      // interior flux points, derivative, accumulation (k_pred)
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
        prims<2>(M, W, a, gam);
        flux<2>(W, F[fp], a, gam);
      }
      const B rha = (B)g.rh[a];
      const int lb = (a == 0) ? lp * n : lp;
      const int sa = (a == 0) ? 1 : n;
      B* d0 = dudt + (long long)(el * T + t) * SLAB + lb;
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

  // Picard update: out[tn] = U_sp - dt * sum_p invader[tn, p] dudt[p]
  const double dt = *dt_dev;
  for (int idx = tid; idx < L::TE * V * NL; idx += blockDim.x) {
    const int le = idx / (V * NL);
    const int v = (idx / NL) % V;
    const int l = idx % NL;
    const int xe = bx * TX + le % TX, ye = by * TY + le / TX;
    if (xe >= Nx || ye >= Ny) continue;
    const long long e = (long long)ye * Nx + xe;
    const long long off = ((long long)v * NE + e) * NL + l;
    const B* dd = dudt + (long long)le * T * SLAB + v * NL + l;
    if (FIRST && out_cmp) {  // the increment alone (compact first iterate)
      out[off] = dd[0];
      continue;
    }
    const double u0 = Usp[off];
#pragma unroll
    for (int tn = 0; tn < L::T; ++tn) {
      B acc = B(0);
#pragma unroll
      for (int p = 0; p < L::T; ++p) {
        const int ps = (p < Tsrc) ? p : (Tsrc - 1);
        acc += op.inv[tn][p] * dd[(long long)ps * SLAB];
      }
      out[((long long)(tn * V + v) * NE + e) * NL + l] = (B)(u0 - (double)acc * dt);
    }
  }
}
template <int P, int RS, typename S, typename B, bool FIRST, bool SRC_CMP>
__global__ void __launch_bounds__(256)
k_pred2d(Ops<P, B> op, Geom g, Params prm, const S* __restrict__ src, int Tsrc,
         const double* __restrict__ Usp, const double* __restrict__ dt_dev, B* __restrict__ out,
         int out_cmp) {
  k_pred2d_body<P, RS, S, B, FIRST, SRC_CMP>(hw_block(), op, g, prm, src, Tsrc, Usp, dt_dev, out, out_cmp);
}

}  // namespace pred2d
}  // namespace spd
