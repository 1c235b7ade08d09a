// 2-D LDS-tiled kernels of the finite-volume corrector: high-order face
// fluxes, trouble detection, blending coefficients and the MUSCL-Hancock
// correction.  Each block owns a tile of cells (or a strip of elements),
// stages what it needs (tile + halo, ghost layers resolved by the index maps
// of spd_device.hpp) in LDS and computes its outputs from there.  The
// arithmetic is that of the generic kernels in spd_step.hip, which this file
// is included into (it uses trouble_at, ring1, kRules2 and Ops from there).
#pragma once

#include <type_traits>

namespace spd {
namespace fv2d {

constexpr int TX = 32;  // tile width (cells along x)
constexpr int TY = 8;   // tile height (cells along y)
constexpr int NT = TX * TY;

__device__ __forceinline__ double qnan() { return __longlong_as_double(0x7ff8000000000000ll); }

__device__ __forceinline__ bool in_ghost_range(const Geom& g, int x, int y) {
  return x >= -NGH && x < g.Ns[0] + NGH && y >= -NGH && y < g.Ns[1] + NGH;
}

// ---------------------------------------------------------------------------
// Trouble detection (src/trouble_detection.py:5-71) on a DTX x DTY tile
// (512 threads: the taller tile stages fewer halo cells per cell).
//   wo: old primitives on the tile + 1 halo (all variables)
//   wn: new (rho, P) on the tile + 3 halo (smooth-extrema stencil)
// ---------------------------------------------------------------------------
constexpr int DTX = 32, DTY = 16, DNT = DTX * DTY;
constexpr int DH = 3;
constexpr int DNX = DTX + 2 * DH, DNY = DTY + 2 * DH;
constexpr int ONX = DTX + 2, ONY = DTY + 2;

// Conserved state of the ghost-filled array at ghosted cell (x, y) (the
// mapping of load_W); false outside the ghosted grid.  fl: reflective flips of
// the x and y velocities.
__device__ __forceinline__ bool stage_cons(const double* __restrict__ U, const Geom& g, int x, int y,
                                           double* u, bool* fl) {
  if (!in_ghost_range(g, x, y)) return false;
  fl[0] = fl[1] = false;
  const int mx = ghost_map(x, g.Ns[0], g.bc[0], &fl[0]);
  const int my = ghost_map(y, g.Ns[1], g.bc[1], &fl[1]);
  const long long idx = (long long)my * g.Ns[0] + mx;
#pragma unroll
  for (int v = 0; v < 4; ++v) u[v] = U[v * g.Nc + idx];
  return true;
}

// Smooth-extrema test along one axis from LDS (src/trouble_detection.py:
// 104-137).  w points at the cell of coordinate x; w[k*stride] is its k-th
// neighbour along the axis.  dU is zero outside [-1, Ns].
__device__ __forceinline__ double sed_dU(const double* w, int stride, int x, int Ns, const double* rdc) {
  return (x >= -1 && x <= Ns) ? (w[stride] - w[-stride]) * rdc[x] : 0.0;
}

// The reference forms alpha_L = -c_L / dv and alpha_R = c_R / dv (c the part
// of v_L, v_R of the sign that limits; 1 where dv == 0; clipped to 1) and
// keeps a candidate flag where the neighbourhood minimum of
// A = min(alpha_R, alpha_L) is below 1.  Every clipped quotient is |c| / |dv|
// in [0, 1], and a correctly rounded |c| / |dv| is below 1 exactly when
// |c| < |dv| (NaN and dv == 0 give 1 there and false here), so the test needs
// no division.  Returns bit 0: A < 1 (true outside [0, Ns), where the
// reference's A is 0), bit 1: alpha_L < 1.  sed_bits_core takes the first
// differences at x - 1, x, x + 1 and rdc, hf at x.
__device__ __forceinline__ int sed_bits_core(double dUm, double dU0, double dUp, double rdc_x,
                                             double hf_x) {
  const double d2U = (dUp - dUm) * rdc_x;
  const double dv = 0.5 * hf_x * d2U;
  const double vL = dUm - dU0;
  const double vR = dUp - dU0;
  const double cL = (dv < 0.0) ? ((vL > 0.0) ? vL : 0.0) : ((vL < 0.0) ? vL : 0.0);
  const double cR = (dv > 0.0) ? ((vR > 0.0) ? vR : 0.0) : ((vR < 0.0) ? vR : 0.0);
  const double adv = fabs(dv);
  const bool l = fabs(cL) < adv, r = fabs(cR) < adv;
  return ((l || r) ? 1 : 0) | (l ? 2 : 0);
}
__device__ __forceinline__ int sed_bits(const double* w, int stride, int x, int Ns, const double* rdc,
                                        const double* hf) {
  if (x < 0 || x >= Ns) return 1;
  const double dUm = sed_dU(w - stride, stride, x - 1, Ns, rdc);
  const double dU0 = sed_dU(w, stride, x, Ns, rdc);
  const double dUp = sed_dU(w + stride, stride, x + 1, Ns, rdc);
  return sed_bits_core(dUm, dU0, dUp, rdc[x], hf[x]);
}

// The same with rdc and hf from a block's LDS copy: rl, hl hold them at the
// local index xl of coordinate x.
__device__ __forceinline__ double sed_dU_l(const double* w, int stride, int x, int Ns,
                                          const double* rl, int xl) {
  return (x >= -1 && x <= Ns) ? (w[stride] - w[-stride]) * rl[xl] : 0.0;
}
__device__ __forceinline__ int sed_bits_l(const double* w, int stride, int x, int Ns,
                                          const double* rl, const double* hl, int xl) {
  if (x < 0 || x >= Ns) return 1;
  const double dUm = sed_dU_l(w - stride, stride, x - 1, Ns, rl, xl - 1);
  const double dU0 = sed_dU_l(w, stride, x, Ns, rl, xl);
  const double dUp = sed_dU_l(w + stride, stride, x + 1, Ns, rl, xl + 1);
  return sed_bits_core(dUm, dU0, dUp, rl[xl], hl[xl]);
}

// alpha < 1 along one axis for the cell at ca, with the compute_min edge rules
// (first cell: min(A[0], A[1]); last cell: min(A[-2], alpha_L[-1])).
__device__ __forceinline__ bool sed_relax_lds(const double* w, int stride, int ca, int Ns,
                                              const double* rdc, const double* hf) {
  const int am = sed_bits(w - stride, stride, ca - 1, Ns, rdc, hf);
  const int a0 = sed_bits(w, stride, ca, Ns, rdc, hf);
  const int ap = sed_bits(w + stride, stride, ca + 1, Ns, rdc, hf);
  if (Ns == 1) return (a0 & 2) != 0;
  if (ca == 0) return ((a0 | ap) & 1) != 0;
  if (ca == Ns - 1) return ((am & 1) | ((a0 >> 1) & 1)) != 0;
  return ((am | a0 | ap) & 1) != 0;
}

// LDS of k_detect2d: the old primitives (tile + 1) are needed by the NAD phase
// only; their space then holds the smooth-extrema bits of the tile (+1 along
// the axis) for rho and P.
constexpr int DWO = 4 * ONY * ONX;
constexpr int DAX = 2 * DTY * (DTX + 2), DAY = 2 * (DTY + 2) * DTX;
static_assert(DAX + DAY <= DWO * 8, "smooth-extrema table must fit the old-state tile");

// New density and pressure of the DTX x DTY tile at (x0, y0) and its DH halo
// (ghost layers by the index maps, stored ghosts of "ic" / "pressure" sides;
// NaN outside the ghosted grid).  Each thread issues the loads of all its
// cells before it uses the first.
__device__ __forceinline__ void stage_new_rho_p(const Geom& g, const Params& prm,
                                                const double* __restrict__ Unew, int x0, int y0,
                                                double (*wn)[DNY][DNX]) {
  constexpr int NI = (DNX * DNY + DNT - 1) / DNT;
  double u[NI][4];
  bool ok[NI], fl[NI][2];
#pragma unroll
  for (int j = 0; j < NI; ++j) {
    const int k = threadIdx.x + j * DNT;
    ok[j] = k < DNX * DNY && stage_cons(Unew, g, x0 + k % DNX - DH, y0 + k / DNX - DH, u[j], fl[j]);
  }
#pragma unroll
  for (int j = 0; j < NI; ++j) {
    const int k = threadIdx.x + j * DNT;
    if (k >= DNX * DNY) continue;
    double r = qnan(), p = qnan();
    if (ok[j]) {
      double W[4];
      prims<2>(u[j], W, 0, prm.gamma);
      int c[3] = {x0 + k % DNX - DH, y0 + k / DNX - DH, 0};
      stored_ghost<2>(g, c, W);  // ghosts of "ic" / "pressure" sides
      r = W[0];
      p = W[3];
    }
    wn[0][k / DNX][k % DNX] = r;
    wn[1][k / DNX][k % DNX] = p;
  }
}

// Subcell geometry of a DTX x DTY tile at (x0, y0) in LDS: rdc and hf along
// x at x = x0 - DH + j (j < DNX) and along y at y = y0 - DH + j (j < DNY);
// zero where the smooth-extrema test does not read them.
struct SedGeo {
  double rdcx[DNX], hfx[DNX], rdcy[DNY], hfy[DNY];
};
__device__ __forceinline__ void stage_sed_geo(const Geom& g, int x0, int y0, SedGeo* sg) {
  const int Nx = g.Ns[0], Ny = g.Ns[1];
  for (int j = threadIdx.x; j < 2 * DNX + 2 * DNY; j += DNT) {
    if (j < DNX) {
      const int x = x0 - DH + j;
      sg->rdcx[j] = (x >= -1 && x <= Nx) ? g.rdc[0][x + NGH] : 0.0;
    } else if (j < 2 * DNX) {
      const int x = x0 - DH + (j - DNX);
      sg->hfx[j - DNX] = (x >= 0 && x < Nx) ? g.hfp[0][x + NGH] : 0.0;
    } else if (j < 2 * DNX + DNY) {
      const int y = y0 - DH + (j - 2 * DNX);
      sg->rdcy[j - 2 * DNX] = (y >= -1 && y <= Ny) ? g.rdc[1][y + NGH] : 0.0;
    } else {
      const int y = y0 - DH + (j - 2 * DNX - DNY);
      sg->hfy[j - 2 * DNX - DNY] = (y >= 0 && y < Ny) ? g.hfp[1][y + NGH] : 0.0;
    }
  }
}

// Smooth-extrema relaxation (src/trouble_detection.py:38-51,104-137) of the
// density / pressure candidates pt[0], pt[3] of the calling thread's cell
// (thread (tx, ty) of the DTX x DTY tile at (x0, y0)): the tile's bits table
// (DAX + DAY bytes of LDS at Ax) is built from wn, the new density and
// pressure on the tile + DH halo, and a candidate is kept where
// alpha = min(1, alpha_x, alpha_y) < 1.  With sg (the tile's geometry in
// LDS, stage_sed_geo) each thread builds the entries of its own column and
// row of the table.  Every thread of the block calls it.
__device__ __forceinline__ void sed_relax_tile(const Geom& g, double (*wn)[DNY][DNX],
                                               uint8_t* Ax, int x0, int y0, bool inside,
                                               bool* pt, const SedGeo* sg = nullptr) {
  const int Nx = g.Ns[0], Ny = g.Ns[1];
  const int tx = threadIdx.x % DTX, ty = threadIdx.x / DTX;
  const int x = x0 + tx, y = y0 + ty;
  uint8_t* Ay = Ax + DAX;  // Ax: [2][DTY][DTX+2], x in [x0-1, x0+DTX]; Ay: [2][DTY+2][DTX]
  const double* rdcx = g.rdc[0] + NGH;
  const double* rdcy = g.rdc[1] + NGH;
  const double* hfx = g.hfp[0] + NGH;
  const double* hfy = g.hfp[1] + NGH;
  if (sg) {
#pragma unroll
    for (int var = 0; var < 2; ++var) {
      // x bits of row ty at lx = tx, tx + DTX (x = x0 + lx - 1)
      for (int lx = tx; lx < DTX + 2; lx += DTX) {
        const int xl = lx - 1 + DH;
        Ax[(var * DTY + ty) * (DTX + 2) + lx] = (uint8_t)sed_bits_l(
            &wn[var][ty + DH][xl], 1, x0 + lx - 1, Nx, sg->rdcx, sg->hfx, xl);
      }
      // y bits of column tx at ly = ty, ty + DTY (y = y0 + ly - 1)
      for (int ly = ty; ly < DTY + 2; ly += DTY) {
        const int yl = ly - 1 + DH;
        Ay[(var * (DTY + 2) + ly) * DTX + tx] = (uint8_t)sed_bits_l(
            &wn[var][yl][tx + DH], DNX, y0 + ly - 1, Ny, sg->rdcy, sg->hfy, yl);
      }
    }
  } else {
    for (int k = threadIdx.x; k < DAX + DAY; k += DNT) {
      if (k < DAX) {
        const int var = k / (DTY * (DTX + 2)), r = k % (DTY * (DTX + 2));
        const int ly = r / (DTX + 2), lx = r % (DTX + 2);
        const double* w = &wn[var][ly + DH][lx - 1 + DH];
        Ax[k] = (uint8_t)sed_bits(w, 1, x0 + lx - 1, Nx, rdcx, hfx);
      } else {
        const int kk = k - DAX;
        const int var = kk / ((DTY + 2) * DTX), r = kk % ((DTY + 2) * DTX);
        const int ly = r / DTX, lx = r % DTX;
        const double* w = &wn[var][ly - 1 + DH][lx + DH];
        Ay[kk] = (uint8_t)sed_bits(w, DNX, y0 + ly - 1, Ny, rdcy, hfy);
      }
    }
  }
  __syncthreads();
  if (!inside) return;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int var = (k == 0) ? 0 : 3;
    if (!pt[var]) continue;
    const double* w = &wn[k][ty + DH][tx + DH];
    const uint8_t* ax = Ax + (k * DTY + ty) * (DTX + 2) + tx + 1;  // bits at x
    const uint8_t* ay = Ay + (k * (DTY + 2) + ty + 1) * DTX + tx;  // bits at y
    bool r1, r2;
    if (Nx == 1 || x == Nx - 1)
      r1 = sed_relax_lds(w, 1, x, Nx, rdcx, hfx);
    else if (x == 0)
      r1 = ((ax[0] | ax[1]) & 1) != 0;
    else
      r1 = ((ax[-1] | ax[0] | ax[1]) & 1) != 0;
    if (Ny == 1 || y == Ny - 1)
      r2 = sed_relax_lds(w, DNX, y, Ny, rdcy, hfy);
    else if (y == 0)
      r2 = ((ay[0] | ay[DTX]) & 1) != 0;
    else
      r2 = ((ay[-DTX] | ay[0] | ay[DTX]) & 1) != 0;
    // alpha = min(1, alpha_x, alpha_y) < 1
    pt[var] = pt[var] && (r1 || r2);
  }
}

// FP64 in both precision forms: the smooth-extrema test takes second
// differences of the new state, ~1e-8 of it at 4096 subcells per side.
__global__ void __launch_bounds__(DNT)
k_detect2d(Geom g, Params prm, const double* __restrict__ Ucv, const double* __restrict__ Unew,
           uint8_t* __restrict__ flags, int* __restrict__ any_flag) {
  SPD_SKIP(g);
  __shared__ double wn[2][DNY][DNX];  // new rho, P
  __shared__ double wbuf[DWO];        // old primitives, then the smooth-extrema bits
  double(*wo)[ONY][ONX] = reinterpret_cast<double(*)[ONY][ONX]>(wbuf);
  int bx, by;
  xcd_band_tile(bx, by);
  const int x0 = bx * DTX, y0 = by * DTY;
  const int Nx = g.Ns[0], Ny = g.Ns[1];
  // staging: each thread issues the loads of all its cells of a tile before
  // it uses the first (new rho, P on tile + 3, then old primitives on tile + 1)
  stage_new_rho_p(g, prm, Unew, x0, y0, wn);
  {
    constexpr int NI = (ONX * ONY + DNT - 1) / DNT;
    double u[NI][4];
    bool ok[NI], fl[NI][2];
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const int k = threadIdx.x + j * DNT;
      ok[j] = k < ONX * ONY && stage_cons(Ucv, g, x0 + k % ONX - 1, y0 + k / ONX - 1, u[j], fl[j]);
    }
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const int k = threadIdx.x + j * DNT;
      if (k >= ONX * ONY) continue;
      double W[4] = {qnan(), qnan(), qnan(), qnan()};
      if (ok[j]) {
        prims<2>(u[j], W, 0, prm.gamma);
        if (fl[j][0]) W[1] = -W[1];
        if (fl[j][1]) W[2] = -W[2];
        int c[3] = {x0 + k % ONX - 1, y0 + k / ONX - 1, 0};
        stored_ghost<2>(g, c, W);  // ghosts of "ic" / "pressure" sides
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) wo[v][k / ONX][k % ONX] = W[v];
    }
  }
  __syncthreads();
  const int tx = threadIdx.x % DTX, ty = threadIdx.x / DTX;
  const int x = x0 + tx, y = y0 + ty;
  const bool inside = x < Nx && y < Ny;
  bool cand = false;
  bool pt[4] = {false, false, false, false};
  double Wn[4];
  if (inside) {
    const int ox = tx + 1, oy = ty + 1;
    Wn[0] = wn[0][ty + DH][tx + DH];
    Wn[3] = wn[1][ty + DH][tx + DH];
    Wn[1] = Wn[2] = 0.0;
    const int nvar_nad = any_flag ? 4 : 2;  // velocities only feed the global gate
    if (any_flag) {
      int c[3] = {x, y, 0};
      double W[4];
      load_W<2>(Unew, g, c, W, prm.gamma);
      Wn[1] = W[1];
      Wn[2] = W[2];
    }
    const double tol = prm.tol;
    for (int k = 0; k < nvar_nad; ++k) {
      const int v = (k == 0) ? 0 : (k == 1 ? 3 : k - 1);
      double mx = wo[v][oy][ox], mn = mx;
      const double nb[4] = {wo[v][oy][ox - 1], wo[v][oy][ox + 1], wo[v][oy - 1][ox],
                            wo[v][oy + 1][ox]};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        mx = nmax(mx, nb[q]);
        mn = nmin(mn, nb[q]);
      }
      double lo = mn, hi = mx;
      if (prm.p > 0) {
        if (prm.nad_delta) {
          const double eps = tol * (hi - lo);
          lo = lo - eps;
          hi = hi + eps;
        } else {
          lo = lo - fabs(lo) * tol;
          hi = hi + fabs(hi) * tol;
        }
      }
      pt[v] = !(Wn[v] >= lo) || !(Wn[v] <= hi);
      cand = cand || pt[v];
    }
  }
  if (prm.p > 1 && prm.sed) {
    // smooth-extrema relaxation of the rho / P candidates; the table is built
    // only for tiles that hold one
    const int need = __syncthreads_or((pt[0] || pt[3]) ? 1 : 0);
    if (need) sed_relax_tile(g, wn, reinterpret_cast<uint8_t*>(wbuf), x0, y0, inside, pt);
  }
  if (inside) {
    bool tr = pt[0] || pt[3];
    if (prm.pad) {
      if (!(Wn[0] >= prm.min_rho) || !(Wn[0] <= prm.max_rho) || !(Wn[3] >= prm.min_P)) tr = true;
    }
    flags[(long long)y * Nx + x] = tr ? 1 : 0;
  }
  if (any_flag) {
    const int blk = __syncthreads_or(cand ? 1 : 0);
    if (blk && threadIdx.x == 0 && *(volatile int*)any_flag == 0) atomicOr(any_flag, 1);
  }
}

// ---------------------------------------------------------------------------
// Second half of the detection when the high-order flux kernel has run its
// first half (k_hocand2d with DET; src/trouble_detection.py:38-55): that
// kernel leaves per cell the code PAD (bit 0) | NAD density candidate (bit 1)
// | NAD pressure candidate (bit 2).  A DTX x DTY tile without candidates
// returns after one vote (its codes are already the final 0 / 1); a tile with
// one stages the new density and pressure on the tile + 3 halo, applies the
// smooth-extrema relaxation of k_detect2d to its candidates and writes their
// final flags.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void k_sed2d_body(const BlockIx bi, Geom g, Params prm, const double* __restrict__ Unew, uint8_t* __restrict__ flags) {
  SPD_SKIP(g);
  __shared__ double wn[2][DNY][DNX];
  __shared__ SedGeo sg;
  __shared__ uint8_t ab[DAX + DAY];
  int bx, by;
  xcd_band_tile(bi, bx, by);
  const int x0 = bx * DTX, y0 = by * DTY;
  const int Nx = g.Ns[0], Ny = g.Ns[1];
  const int x = x0 + threadIdx.x % DTX, y = y0 + threadIdx.x / DTX;
  const bool inside = x < Nx && y < Ny;
  const int code = inside ? (int)flags[(long long)y * Nx + x] : 0;
  if (!__syncthreads_or(code & 6)) return;
  stage_sed_geo(g, x0, y0, &sg);
  stage_new_rho_p(g, prm, Unew, x0, y0, wn);
  __syncthreads();
  bool pt[4] = {(code & 2) != 0, false, false, (code & 4) != 0};
  sed_relax_tile(g, wn, ab, x0, y0, inside, pt, &sg);
  if (inside && (code & 6)) flags[(long long)y * Nx + x] = (pt[0] || pt[3] || (code & 1)) ? 1 : 0;
}
__global__ void __launch_bounds__(DNT)
k_sed2d(Geom g, Params prm, const double* __restrict__ Unew, uint8_t* __restrict__ flags) {
  k_sed2d_body(hw_block(), g, prm, Unew, flags);
}

// ---------------------------------------------------------------------------
// Blending coefficients theta on the ghosted grid (src/trouble_detection.py:
// 73-87,139-169): trouble values on the tile + 2 halo, theta1 on tile + 1
// halo, then the x and y 0.25 rings.  Tiles cover the ghosted grid; (gx, gy)
// are ghosted indices.  Without blending theta = trouble.  Tiles of active
// cells take the byte form below (flags only, no FP64 work); tiles that reach
// the ghost layers take the general form.
// ---------------------------------------------------------------------------
// 64 x 16 outputs per block of 256 threads (four per thread).
constexpr int HTX = 64, HTY = 16, HNT = 256;
constexpr int TH = 2;
constexpr int TNX = HTX + 2 * TH, TNY = HTY + 2 * TH;

// Tiles of the compacted fallback that a block of k_theta2d marks: with
// `marks` set, every cell whose theta is not 0 marks the fallback tiles
// (cells of mk.tx x mk.ty, mk.ntx per row) holding it and its four
// neighbours (the cells that border its faces), through a flag table of the
// block in LDS; the list kernel reads and clears the marks.
struct FbMarks {
  uint8_t* marks;
  int tx, ty, ntx, nty;
  // per theta tile (HTX x HTY, row-major over the launch grid): 1 where the
  // tile may hold non-zero theta; a tile of the byte form without troubled
  // cells that is not marked writes nothing (its theta is 0 already).  Null:
  // every tile writes.
  uint8_t* dirty = nullptr;
};
constexpr int MKT = 8;  // flag table of a block: MKT x MKT fallback tiles

__device__ __forceinline__ void k_theta2d_body(const BlockIx bi, Geom g, Params prm, const uint8_t* __restrict__ flags, const double* __restrict__ Ucv,
          const double* __restrict__ Unew, const int* __restrict__ any_flag,
          double* __restrict__ theta, FbMarks mk) {
  SPD_SKIP(g);
  __shared__ double tr[TNY][TNX];
  __shared__ double t1[HTY + 2][HTX + 2];
  // fallback-tile marks: the table bits (two words), and the table column /
  // row of the cells x = gx0 - NGH - 1 + j, j < HTX + 2, and
  // y = gy0 - NGH - 1 + j, j < HTY + 2 (-1 outside the grid or the table)
  __shared__ unsigned smk[2];
  __shared__ signed char mcol[HTX + 2], mrow[HTY + 2];
  const int gx0 = bi.x * HTX, gy0 = bi.y * HTY;
  const int NGx = g.Ng[0], NGy = g.Ng[1];
  // first fallback tile of the block's table (cells from gx0 - NGH - 1)
  const int mbx = mk.marks ? ((gx0 - NGH - 1 > 0) ? gx0 - NGH - 1 : 0) / mk.tx : 0;
  const int mby = mk.marks ? ((gy0 - NGH - 1 > 0) ? gy0 - NGH - 1 : 0) / mk.ty : 0;
  if (mk.marks) {
    const int t = threadIdx.x;
    if (t < 2) smk[t] = 0;
    if (t < HTX + 2) {
      const int x = gx0 - NGH - 1 + t;
      const int c = (x >= 0 && x < g.Ns[0]) ? x / mk.tx - mbx : -1;
      mcol[t] = (signed char)((c < MKT) ? c : -1);
    } else if (t < HTX + 2 + HTY + 2) {
      const int y = gy0 - NGH - 1 + (t - HTX - 2);
      const int r = (y >= 0 && y < g.Ns[1]) ? y / mk.ty - mby : -1;
      mrow[t - HTX - 2] = (signed char)((r < MKT) ? r : -1);
    }
  }
  // theta of ghosted cell (gx, gy) of the block's outputs is not 0: the
  // table bits of the tiles of the cell and of its four neighbours
  unsigned long long mbits = 0;
  auto mark = [&](int gx, int gy) {
    const int jx = gx - gx0 + 1, jy = gy - gy0 + 1;
    const int cj[5] = {mcol[jx], mcol[jx - 1], mcol[jx + 1], mcol[jx], mcol[jx]};
    const int rj[5] = {mrow[jy], mrow[jy], mrow[jy], mrow[jy - 1], mrow[jy + 1]};
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (cj[q] >= 0 && rj[q] >= 0) mbits |= 1ull << (rj[q] * MKT + cj[q]);
  };
  // after the block's last barrier: write the flagged tiles' marks
  auto flush_marks = [&]() {
    // OR over the wave first: one LDS atomic per wave and word
    unsigned lo = (unsigned)mbits, hi = (unsigned)(mbits >> 32);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo |= __shfl_xor(lo, off);
      hi |= __shfl_xor(hi, off);
    }
    if ((threadIdx.x & 63) == 0) {
      if (lo) atomicOr(&smk[0], lo);
      if (hi) atomicOr(&smk[1], hi);
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < MKT * MKT && ((smk[t / 32] >> (t % 32)) & 1u)) {
      const int tx = mbx + t % MKT, ty = mby + t / MKT;
      if (tx < mk.ntx && ty < mk.nty) mk.marks[(long long)ty * mk.ntx + tx] = 1;
    }
  };
  // A tile whose staged cells are all active cells (most tiles) holds trouble
  // values 0 / 1 only, and with blending theta is then a function of the
  // distance to the nearest troubled cell (the 3 x 3 maximum of theta1 with
  // weights 1, 0.75 (edge neighbour), 0.5 (corner), followed by the two 0.25
  // rings, which set 0.25 wherever theta1 = 0 has a non-zero theta1 in its
  // 3 x 3 neighbourhood):
  //   1 troubled, 0.75 edge neighbour of one, 0.5 corner neighbour of one,
  //   0.25 a troubled cell within the 5 x 5 neighbourhood, else 0.
  // That case runs on bytes: flags, then per staged row a code of the 1-, 3-
  // and 5-wide horizontal ORs, then the rows' codes are combined per output.
  const int cx0 = gx0 - TH - NGH, cy0 = gy0 - TH - NGH;  // cell coordinates of tr[0][0]
  if (prm.blending && cx0 >= 0 && cy0 >= 0 && cx0 + TNX <= g.Ns[0] && cy0 + TNY <= g.Ns[1]) {
    // read by every thread before the barrier below, cleared or set after it
    const long long tile = (long long)bi.y * ((NGx + HTX - 1) / HTX) + bi.x;
    const bool was_dirty = mk.dirty ? mk.dirty[tile] != 0 : true;
    uint8_t* fb = reinterpret_cast<uint8_t*>(&tr[0][0]);  // [TNY][TNX] flags
    uint8_t* hc = reinterpret_cast<uint8_t*>(&t1[0][0]);  // [TNY][HTX] row codes
    constexpr int NIT = (TNX * TNY + HNT - 1) / HNT;
    int f[NIT];
    int any = 0;
#pragma unroll
    for (int j = 0; j < NIT; ++j) {  // every load issued before the first is used
      const int k = threadIdx.x + j * HNT;
      f[j] = (k < TNX * TNY)
                 ? (int)flags[(long long)(cy0 + k / TNX) * g.Ns[0] + cx0 + k % TNX]
                 : 0;
    }
#pragma unroll
    for (int j = 0; j < NIT; ++j) {
      const int k = threadIdx.x + j * HNT;
      if (k < TNX * TNY) fb[k] = (uint8_t)f[j];
      any |= f[j];
    }
    if (!__syncthreads_or(any)) {
      if (!was_dirty) return;
      for (int k = threadIdx.x; k < HTX * HTY; k += HNT)
        theta[(long long)(gy0 + k / HTX) * NGx + gx0 + k % HTX] = 0.0;
      if (mk.dirty && threadIdx.x == 0) mk.dirty[tile] = 0;
      return;
    }
    if (mk.dirty && !was_dirty && threadIdx.x == 0) mk.dirty[tile] = 1;
    for (int k = threadIdx.x; k < TNY * HTX; k += HNT) {
      const uint8_t* r = fb + (k / HTX) * TNX + k % HTX + TH;
      const int c = r[0], h1 = r[-1] | c | r[1], h2 = r[-2] | h1 | r[2];
      hc[k] = (uint8_t)(c | (h1 << 1) | (h2 << 2));
    }
    __syncthreads();
    for (int k = threadIdx.x; k < HTX * HTY; k += HNT) {
      const uint8_t* col = hc + (k / HTX + TH) * HTX + k % HTX;
      const int a = col[0], u1 = col[-HTX], d1 = col[HTX];
      const int r2 = (a | u1 | d1 | col[-2 * HTX] | col[2 * HTX]) >> 2;
      double th = 0.0;
      if (a & 1)
        th = 1.0;
      else if (((a >> 1) | u1 | d1) & 1)
        th = 0.75;
      else if (((u1 | d1) >> 1) & 1)
        th = 0.5;
      else if (r2 & 1)
        th = 0.25;
      theta[(long long)(gy0 + k / HTX) * NGx + gx0 + k % HTX] = th;
      if (mk.marks && th != 0.0) mark(gx0 + k % HTX, gy0 + k / HTX);
    }
    if (mk.marks) flush_marks();
    return;
  }
  {
    const bool gate = (*any_flag != 0) && prm.p > 1 && prm.sed;
    const double* Ulast = gate ? Unew : Ucv;
    for (int k = threadIdx.x; k < TNX * TNY; k += HNT) {
      const int lx = k % TNX, ly = k / TNX;
      const int gx = gx0 + lx - TH, gy = gy0 + ly - TH;
      double v = 0.0;
      if (gx >= 0 && gx < NGx && gy >= 0 && gy < NGy) {
        int c[3] = {gx - NGH, gy - NGH, 0};
        v = trouble_at<2>(g, flags, Ulast, c);
      }
      tr[ly][lx] = v;
    }
  }
  __syncthreads();
  if (!prm.blending) {
    for (int k = threadIdx.x; k < HTX * HTY; k += HNT) {
      const int tx = k % HTX, ty = k / HTX;
      const int gx = gx0 + tx, gy = gy0 + ty;
      if (gx < NGx && gy < NGy) {
        const double th = tr[ty + TH][tx + TH];
        theta[(long long)gy * NGx + gx] = th;
        if (mk.marks && !(th == 0.0)) mark(gx, gy);
      }
    }
    if (mk.marks) flush_marks();
    return;
  }
  for (int k = threadIdx.x; k < (HTX + 2) * (HTY + 2); k += HNT) {
    const int lx = k % (HTX + 2), ly = k / (HTX + 2);
    const int gx = gx0 + lx - 1, gy = gy0 + ly - 1;
    double th = 0.0;
    if (gx >= 0 && gx < NGx && gy >= 0 && gy < NGy) {
      const int sx = lx + 1, sy = ly + 1;  // position in tr
      th = tr[sy][sx];
      if (gx >= 1) th = nmax(0.75 * tr[sy][sx - 1], th);
      if (gx <= NGx - 2) th = nmax(0.75 * tr[sy][sx + 1], th);
      if (gy >= 1) th = nmax(0.75 * tr[sy - 1][sx], th);
      if (gy <= NGy - 2) th = nmax(0.75 * tr[sy + 1][sx], th);
      // diagonals (target slice of each update, src/trouble_detection.py:148-151)
      if (gx <= NGx - 2 && gy <= NGy - 2) th = nmax(0.5 * tr[sy + 1][sx + 1], th);
      if (gx >= 1 && gy <= NGy - 2) th = nmax(0.5 * tr[sy + 1][sx - 1], th);
      if (gx <= NGx - 2 && gy >= 1) th = nmax(0.5 * tr[sy - 1][sx + 1], th);
      if (gx >= 1 && gy >= 1) th = nmax(0.5 * tr[sy - 1][sx - 1], th);
    }
    t1[ly][lx] = th;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < HTX * HTY; k += HNT) {
    const int tx = k % HTX, ty = k / HTX;
    const int gx = gx0 + tx, gy = gy0 + ty;
    if (gx >= NGx || gy >= NGy) continue;
    const int lx = tx + 1, ly = ty + 1;
    auto rx = [&](int dy) -> double {
      const double lo = (gx >= 1) ? t1[ly + dy][lx - 1] : 0.0;
      const double hi = (gx <= NGx - 2) ? t1[ly + dy][lx + 1] : 0.0;
      return ring1(gx, NGx, lo, t1[ly + dy][lx], hi);
    };
    const double lo = (gy >= 1) ? rx(-1) : 0.0;
    const double hi = (gy <= NGy - 2) ? rx(+1) : 0.0;
    const double th = ring1(gy, NGy, lo, rx(0), hi);
    theta[(long long)gy * NGx + gx] = th;
    if (mk.marks && !(th == 0.0)) mark(gx, gy);
  }
  if (mk.marks) flush_marks();
}
__global__ void __launch_bounds__(HNT)
k_theta2d(Geom g, Params prm, const uint8_t* __restrict__ flags, const double* __restrict__ Ucv,
          const double* __restrict__ Unew, const int* __restrict__ any_flag,
          double* __restrict__ theta, FbMarks mk) {
  k_theta2d_body(hw_block(), g, prm, flags, Ucv, Unew, any_flag, theta, mk);
}

// ---------------------------------------------------------------------------
// The fallback's arithmetic per cell and per face (src/muscl.py:12-75,
// 129-221, src/sdader_simulator.py:325-331, src/fv_simulator.py:189-205),
// shared by every 2-D fallback kernel.  Nothing here is contracted into fused
// multiply-adds, so each operation is rounded on its own whatever the kernel
// around it keeps in registers, LDS or other lanes: the column-strip and the
// compacted-tile fallbacks give the same bits, and a substep may take either.
// ---------------------------------------------------------------------------
// Limited slope of one variable along one axis from its two neighbours.
template <typename B>
__device__ __forceinline__ B fb_slope(B Wm, B Wc, B Wp, B rhl, B rhr, B hcl, B hcr, B hf,
                                      int mc) {
#pragma clang fp contract(off)
  const B L = (Wc - Wm) * rhl, R = (Wp - Wc) * rhr;
  return mc ? moncen<B>(L, R, hcl, hcr, hf) : minmod<B>(L, R);
}
// Offset of the face states from the Hancock state: half the slope times the
// cell width.
template <typename B>
__device__ __forceinline__ B fb_half(B dM, B hf) {
#pragma clang fp contract(off)
  return B(0.5) * dM * hf;
}
// Hancock half-step state of a cell from its primitives and limited slopes.
template <typename B>
__device__ __forceinline__ void fb_hancock(const B* Wc, const B* dMx, const B* dMy, B gamma,
                                           B dtib, B* Mp) {
#pragma clang fp contract(off)
  B dtW[4] = {B(0), B(0), B(0), B(0)};
  const B rrho = qrcp<B>(Wc[0]);  // shared by both directions (within an ulp)
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int vel = b + 1;
    const B* dW = (b == 0) ? dMx : dMy;
    dtW[0] = dtW[0] - (Wc[vel] * dW[0] + Wc[0] * dW[vel]);
    dtW[3] = dtW[3] - (Wc[vel] * dW[3] + gamma * Wc[3] * dW[vel]);
    dtW[vel] = dtW[vel] - (Wc[vel] * dW[vel] + dW[3] * rrho);
    dtW[2] = dtW[2] - Wc[vel] * dW[2];
  }
  dtW[3] = dtW[3] * B(0);
#pragma unroll
  for (int v = 0; v < 4; ++v) Mp[v] = Wc[v] + B(0.5) * dtW[v] * dtib;
}
// Blended face flux: theta times the fallback flux plus (1 - theta) times the
// high-order flux.
template <typename B>
__device__ __forceinline__ B fb_blend(B tb, B Ff, B Fh) {
#pragma clang fp contract(off)
  return tb * Ff + (B(1) - tb) * Fh;
}
// Corrected state of a cell from its face fluxes (the expression of k_apply).
__device__ __forceinline__ double fb_apply(double u, double FxL, double FxR, double FyB,
                                           double FyT, double rdx, double rdy, double dti) {
#pragma clang fp contract(off)
  double d = u * 0.0;
  d = d + (FxR - FxL) * rdx;
  d = d + (FyT - FyB) * rdy;
  return u - d * dti;
}

// ---------------------------------------------------------------------------
// MUSCL-Hancock correction on a TX x TY tile of cells (src/muscl.py:61-75,
// 129-221, src/fv_simulator.py:88-175,189-205, src/sdader_simulator.py:325-331):
//   0. theta on tile + 1 halo; a tile none of whose faces has a non-zero
//      coefficient (and no godunov) returns at once: its blended fluxes are
//      the high-order ones, so the candidate state already in U_out is the
//      result (k_apply evaluates the same expression);
//   1. old primitives on tile + 2 halo;
//   2. Hancock state and half slopes of the cells that border a face with a
//      non-zero coefficient (tile + 1 halo, plus-shaped);
//   3. per face of the tile: fallback HLLC/LLF flux where the coefficient is
//      non-zero, blended with the high-order flux;
//   4. U_out = U_cv - dt_i * sum(dF/dx) for the tile's cells.
// U_out holds the candidate state on entry and must differ from U_cv.
// ---------------------------------------------------------------------------
constexpr int CX = TX + 2, CY = TY + 2;   // tile + 1 halo
constexpr int WX = TX + 4, WY = TY + 4;   // tile + 2 halo

template <int RS, typename B>
__global__ void __launch_bounds__(NT)
k_fbcorrect2d(Geom g, Params prm, const double* __restrict__ Ucv,
              const double* __restrict__ theta, const B* __restrict__ Fx,
              const B* __restrict__ Fy, const double* __restrict__ dt_dev, double w_i,
              double* __restrict__ Uout) {
  SPD_SKIP(g);
  // LDS: W (4 x WY x WX) is reused for the blended face fluxes after step 2
  constexpr int SW = (4 * WY * WX > 4 * (TY * (TX + 1) + (TY + 1) * TX))
                         ? 4 * WY * WX
                         : 4 * (TY * (TX + 1) + (TY + 1) * TX);
  __shared__ B sW[SW];
  __shared__ B sH[12][CY][CX];  // Mp(4), Sx(4), Sy(4)
  __shared__ double sT[CY][CX];
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
  const int Nx = g.Ns[0], Ny = g.Ns[1];
  const int NGx = g.Ng[0];
  const B gamma = (B)prm.gamma;
  const double dti = (*dt_dev) * w_i;
  const B dtib = (B)dti;
  B(*W)[WY][WX] = reinterpret_cast<B(*)[WY][WX]>(sW);

  int live = prm.godunov != 0;
  for (int k = threadIdx.x; k < CX * CY; k += NT) {
    const int lx = k % CX, ly = k / CX;
    const int x = x0 + lx - 1, y = y0 + ly - 1;
    double th = 0.0;
    if (x >= -1 && x <= Nx && y >= -1 && y <= Ny)
      th = theta[(long long)(y + NGH) * NGx + (x + NGH)];
    sT[ly][lx] = th;
    live |= !(th == 0.0);
  }
  if (!__syncthreads_or(live)) return;

  for (int k = threadIdx.x; k < WX * WY; k += NT) {
    const int lx = k % WX, ly = k / WX;
    const int x = x0 + lx - 2, y = y0 + ly - 2;
    double Wv[4] = {qnan(), qnan(), qnan(), qnan()};
    if (in_ghost_range(g, x, y)) {
      int c[3] = {x, y, 0};
      load_W<2>(Ucv, g, c, Wv, prm.gamma);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) W[v][ly][lx] = (B)Wv[v];
  }
  __syncthreads();

  // 2. which cells border an active face of this tile
  for (int k = threadIdx.x; k < CX * CY; k += NT) {
    const int lx = k % CX, ly = k / CX;
    const int x = x0 + lx - 1, y = y0 + ly - 1;
    bool need = false;
    const bool godunov = prm.godunov != 0;
    // x-faces of tile rows: cell is L of face x+1 (x in [x0-1, x0+TX-1]) or
    // R of face x (x in [x0, x0+TX]); faces exist for x+1 <= Nx / x <= Nx
    if (ly >= 1 && ly <= TY && y < Ny) {
      if (lx <= TX && x + 1 <= Nx) need |= godunov || !(nmax(sT[ly][lx], sT[ly][lx + 1]) == 0.0);
      if (lx >= 1 && x <= Nx) need |= godunov || !(nmax(sT[ly][lx - 1], sT[ly][lx]) == 0.0);
    }
    if (lx >= 1 && lx <= TX && x < Nx) {
      if (ly <= TY && y + 1 <= Ny) need |= godunov || !(nmax(sT[ly][lx], sT[ly + 1][lx]) == 0.0);
      if (ly >= 1 && y <= Ny) need |= godunov || !(nmax(sT[ly - 1][lx], sT[ly][lx]) == 0.0);
    }
    if (!need) continue;
    const int wx = lx + 1, wy = ly + 1;  // position in W
    B Wc[4], dMh[2][4], Mp[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) Wc[v] = W[v][wy][wx];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int cb = (b == 0) ? x : y;
      const int G = cb + NGH;
      const B rhl = (B)g.rhcv[b][G - 1], rhr = (B)g.rhcv[b][G];
      const B hf = (B)g.hfp[b][G];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const B Wm = (b == 0) ? W[v][wy][wx - 1] : W[v][wy - 1][wx];
        const B Wp = (b == 0) ? W[v][wy][wx + 1] : W[v][wy + 1][wx];
        dMh[b][v] = fb_slope<B>(Wm, Wc[v], Wp, rhl, rhr, (B)g.hcv[b][G - 1], (B)g.hcv[b][G], hf,
                                prm.moncen);
        sH[4 + 4 * b + v][ly][lx] = fb_half<B>(dMh[b][v], hf);
      }
    }
    fb_hancock<B>(Wc, dMh[0], dMh[1], gamma, dtib, Mp);
#pragma unroll
    for (int v = 0; v < 4; ++v) sH[v][ly][lx] = Mp[v];
  }
  __syncthreads();

  // 3. blended fluxes of the tile's faces, in the (now free) W buffer
  B* bFx = sW;                       // [TY][TX+1][4]
  B* bFy = sW + TY * (TX + 1) * 4;   // [TY+1][TX][4]
  const long long nfx = (long long)Ny * (Nx + 1), nfy = (long long)(Ny + 1) * Nx;
  for (int k = threadIdx.x; k < TY * (TX + 1); k += NT) {
    const int fx = k % (TX + 1), ly = k / (TX + 1);
    const int x = x0 + fx, y = y0 + ly;
    if (x > Nx || y >= Ny) continue;
    const long long fi = (long long)y * (Nx + 1) + x;
    B F[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) F[v] = Fx[v * nfx + fi];
    double th = nmax(sT[ly + 1][fx], sT[ly + 1][fx + 1]);
    if (prm.godunov) th = 1.0;
    if (!(th == 0.0)) {
      B WL[4], WR[4], UL[4], UR[4], Ff[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        WL[v] = sH[v][ly + 1][fx] + sH[4 + v][ly + 1][fx];
        WR[v] = sH[v][ly + 1][fx + 1] - sH[4 + v][ly + 1][fx + 1];
      }
      cons<2>(WL, UL, 0, gamma);
      cons<2>(WR, UR, 0, gamma);
      riemann<2, RS>(WL, WR, UL, UR, Ff, 0, gamma, (B)prm.min_c2, prm.lowmach_fv);
      const B tb = (B)th;
#pragma unroll
      for (int v = 0; v < 4; ++v) F[v] = fb_blend<B>(tb, Ff[v], F[v]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) bFx[(ly * (TX + 1) + fx) * 4 + v] = F[v];
  }
  for (int k = threadIdx.x; k < (TY + 1) * TX; k += NT) {
    const int lx = k % TX, fy = k / TX;
    const int x = x0 + lx, y = y0 + fy;
    if (x >= Nx || y > Ny) continue;
    const long long fi = (long long)y * Nx + x;
    B F[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) F[v] = Fy[v * nfy + fi];
    double th = nmax(sT[fy][lx + 1], sT[fy + 1][lx + 1]);
    if (prm.godunov) th = 1.0;
    if (!(th == 0.0)) {
      B WL[4], WR[4], UL[4], UR[4], Ff[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        WL[v] = sH[v][fy][lx + 1] + sH[8 + v][fy][lx + 1];
        WR[v] = sH[v][fy + 1][lx + 1] - sH[8 + v][fy + 1][lx + 1];
      }
      cons<2>(WL, UL, 1, gamma);
      cons<2>(WR, UR, 1, gamma);
      riemann<2, RS>(WL, WR, UL, UR, Ff, 1, gamma, (B)prm.min_c2, prm.lowmach_fv);
      const B tb = (B)th;
#pragma unroll
      for (int v = 0; v < 4; ++v) F[v] = fb_blend<B>(tb, Ff[v], F[v]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) bFy[(fy * TX + lx) * 4 + v] = F[v];
  }
  __syncthreads();

  // 4. apply
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= Nx || y >= Ny) return;
  const double rdx = g.rhfp[0][x + NGH], rdy = g.rhfp[1][y + NGH];
  const long long ci = (long long)y * Nx + x;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    Uout[v * g.Nc + ci] = fb_apply(
        Ucv[v * g.Nc + ci], (double)bFx[(ty * (TX + 1) + tx) * 4 + v],
        (double)bFx[(ty * (TX + 1) + tx + 1) * 4 + v], (double)bFy[(ty * TX + tx) * 4 + v],
        (double)bFy[((ty + 1) * TX + tx) * 4 + v], rdx, rdy, dti);
}

// ---------------------------------------------------------------------------
// The same correction by column strips (the default 2-D form): a wave owns
// MOUT columns of cells (its 64 lanes cover two more on each side) and
// marches up a chunk of MR rows (march_rows), keeping in registers a three-row window of
// old primitives, the Hancock state of the row below and the fluxes the next
// row's update needs; x-neighbours come from the adjacent lanes.  Every
// cell's primitives and Hancock state and every face's fallback flux are
// evaluated once per strip: no tile halo, no LDS, no barriers.  The loads of
// the next row are issued one row ahead.  The arithmetic of each cell and face
// is that of k_fbcorrect2d; only the cells that border a face with a non-zero
// coefficient are written (the others keep the candidate, the same
// expression).
// ---------------------------------------------------------------------------
constexpr int MW = 64, MOUT = MW - 4;  // lanes of a wave, output columns of a strip
constexpr int MWAVES = 4;              // waves (chunks of one strip) per block

// Rows per chunk: 32 (about 8% of the rows evaluated twice, at chunk edges),
// fewer on small grids so that the launch still has some 4 waves per CU.
inline int march_rows(int Nx, int Ny) {
  const long long strips = (Nx + MOUT - 1) / MOUT;
  int r = 32;
  while (r > 4 && strips * ((Ny + r - 1) / r) < 4096) r /= 2;
  return r;
}

template <typename B>
__device__ __forceinline__ B from_right(B v) {  // value of lane + 1 (lane 63: its own)
  return __shfl_down(v, 1, MW);
}
template <typename B>
__device__ __forceinline__ B from_left(B v) {  // value of lane - 1 (lane 0: its own)
  return __shfl_up(v, 1, MW);
}

template <int RS, typename B, bool STORED>
__device__ __forceinline__ void k_fbmarch2d_body(const BlockIx bi, Geom g, Params prm, const double* __restrict__ Ucv,
            const double* __restrict__ theta, const B* __restrict__ Fx,
            const B* __restrict__ Fy, const double* __restrict__ dt_dev, double w_i,
            int MR, double* __restrict__ Uout, const int* __restrict__ gate) {
  SPD_SKIP(g);
  if (gate && *gate == 0) return;  // the adaptive form took the compacted path
  const int lane = threadIdx.x % MW;
  const int Nx = g.Ns[0], Ny = g.Ns[1];
  const int NGx = g.Ng[0];
  const int x = bi.x * MOUT - 2 + lane;
  const int y0 = (bi.y * MWAVES + threadIdx.x / MW) * MR;
  if (y0 >= Ny) return;  // a whole wave
  const int y1 = (y0 + MR < Ny) ? y0 + MR : Ny;
  const B gamma = (B)prm.gamma, mc2 = (B)prm.min_c2;
  const double dti = (*dt_dev) * w_i;
  const B dtib = (B)dti;
  const bool godunov = prm.godunov != 0;
  const bool out = lane >= 2 && lane < MW - 2 && x < Nx;  // lanes that update a cell
  const bool xghost = x >= -NGH && x < Nx + NGH;
  const bool xth = x >= -1 && x <= Nx;                     // theta read (else 0)
  // x geometry of the lane's cell (halo lanes outside [-1, Nx] clamp: unused)
  const int Gx = ((x < -1) ? -1 : (x > Nx ? Nx : x)) + NGH;
  const B rhlx = (B)g.rhcv[0][Gx - 1], rhrx = (B)g.rhcv[0][Gx], hfx = (B)g.hfp[0][Gx];
  const B hclx = (B)g.hcv[0][Gx - 1], hcrx = (B)g.hcv[0][Gx];
  const double rdx = g.rhfp[0][Gx];
  const long long nfx = (long long)Ny * (Nx + 1), nfy = (long long)(Ny + 1) * Nx;
  const bool fxin = x >= -1 && x < Nx;  // the lane's x-face (x + 1/2) exists
  const bool fyin = x >= 0 && x < Nx;   // the lane's y-faces exist

  // conservative state of the ghosted cell (x, y), fetched ahead of its use
  double un[4];
  bool okn = false, flxn = false, flyn = false;
  auto fetch = [&](int y) {
    okn = xghost && y >= -NGH && y < Ny + NGH;
    flxn = flyn = false;
    if (okn) {
      const int mx = ghost_map(x, Nx, g.bc[0], &flxn);
      const int my = ghost_map(y, Ny, g.bc[1], &flyn);
      const long long idx = (long long)my * Nx + mx;
#pragma unroll
      for (int v = 0; v < 4; ++v) un[v] = Ucv[v * g.Nc + idx];
    }
  };
  auto to_W = [&](B* W, int y) {
    double Wv[4] = {qnan(), qnan(), qnan(), qnan()};
    if (okn) {
      prims_rn<2>(un, Wv, 0, prm.gamma);
      if (flxn) Wv[1] = -Wv[1];
      if (flyn) Wv[2] = -Wv[2];
      if constexpr (STORED) {
        int c[3] = {x, y, 0};
        stored_ghost<2>(g, c, Wv);  // ghosts of "ic" / "pressure" sides
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) W[v] = (B)Wv[v];
  };
  auto theta_at = [&](int y) -> double {
    return (xth && y >= -1 && y <= Ny) ? theta[(long long)(y + NGH) * NGx + (x + NGH)] : 0.0;
  };

  // Rows with a non-zero theta anywhere across the wave's columns: bit r for
  // row y0 - 3 + r.  Row k of the march is quiet when rows k - 2 .. k + 2 all
  // have theta = 0: then no face of row k or of its neighbours k - 1, k + 1 is
  // blended, so row k's Hancock state is never used, cell row k - 1 keeps the
  // candidate and the fluxes of row k are not needed.  Quiet rows are
  // skipped; the row window is reloaded at the next row that is not.
  // All rows' theta are loaded at once (the march's registers are not live
  // yet); chunks of more than 34 rows are never skipped.
  constexpr int NBR = 40;
  unsigned long long busy = ~0ull;
  const int nr = y1 - y0 + 6;
  if (!godunov && nr <= NBR) {
    double tv[NBR];
#pragma unroll
    for (int r = 0; r < NBR; ++r) tv[r] = (r < nr) ? theta_at(y0 - 3 + r) : 0.0;
    unsigned long long lb = 0;  // the lane's rows, then OR over the wave
#pragma unroll
    for (int r = 0; r < NBR; ++r)
      if (!(tv[r] == 0.0)) lb |= 1ull << r;
#pragma unroll
    for (int off = MW / 2; off > 0; off >>= 1) lb |= __shfl_xor(lb, off, MW);
    busy = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(lb >> 32)) << 32) |
           (unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)lb);
    if (busy == 0) return;  // the candidate stands everywhere in the chunk
  }
  auto quiet = [&](int k) { return ((busy >> (k - y0 + 1)) & 31ull) == 0; };

  // SKIP: the loop with the quiet-row tests, for chunks that have quiet rows;
  // chunks without one run the plain loop
  auto march = [&](auto skip) {
    constexpr bool SKIP = decltype(skip)::value;
    B Wm[4], W0[4], Wp[4];
    fetch(y0 - 2);
    to_W(Wm, y0 - 2);
    fetch(y0 - 1);
    to_W(W0, y0 - 1);
    fetch(y0);
    double th = theta_at(y0 - 1), thn = theta_at(y0);
    double thm = 0.0;                  // theta of row k - 1
    double tfxR = 0.0, tfyB = 0.0;     // coefficients of FxR and FyB
    B Qm[4] = {B(0), B(0), B(0), B(0)};  // Hancock state + half y-slope of row k - 1
    B FxR[4] = {B(0), B(0), B(0), B(0)};  // blended flux at (k - 1, x + 1/2)
    B FyB[4] = {B(0), B(0), B(0), B(0)};  // blended flux at y-face k - 1
    bool primed = true;  // Wm, W0, un, th, thn hold rows k - 1, k, k + 1
    for (int k = y0 - 1; k <= y1; ++k) {
      if (SKIP && quiet(k)) {
        // theta = 0 on rows k - 2 .. k + 2: nothing of rows k - 1 .. k is
        // blended, and the state of the next busy row is reloaded
        primed = false;
        thm = 0.0;
        tfxR = 0.0;
        tfyB = 0.0;
        continue;
      }
      if (SKIP && !primed) {
        fetch(k - 1);
        to_W(Wm, k - 1);
        fetch(k);
        to_W(W0, k);
        fetch(k + 1);
        th = theta_at(k);
        thn = theta_at(k + 1);
        primed = true;
      }
      to_W(Wp, k + 1);  // row k + 1
      fetch(k + 2);
      const double thk = th;
      th = thn;
      thn = theta_at(k + 2);
      // high-order fluxes this row needs: y-face k, x-face (k, x + 1/2)
      B Fyk[4], Fxk[4];
  #pragma unroll
      for (int v = 0; v < 4; ++v) {
        Fyk[v] = (k >= y0 && fyin) ? Fy[v * nfy + (long long)k * Nx + x] : B(0);
        Fxk[v] = (k >= y0 && k < y1 && fxin) ? Fx[v * nfx + (long long)k * (Nx + 1) + x + 1] : B(0);
      }

      // Hancock state and half slopes of row k (as k_fbcorrect2d step 2)
      const int Gy = k + NGH;  // k in [-1, Ny]
      const B rhly = (B)g.rhcv[1][Gy - 1], rhry = (B)g.rhcv[1][Gy], hfy = (B)g.hfp[1][Gy];
      B Mp[4], Sx[4], Sy[4], dMx[4], dMy[4];
      const B hcly = (B)g.hcv[1][Gy - 1], hcry = (B)g.hcv[1][Gy];
  #pragma unroll
      for (int v = 0; v < 4; ++v) {
        const B Wl = from_left(W0[v]), Wr = from_right(W0[v]);
        dMx[v] = fb_slope<B>(Wl, W0[v], Wr, rhlx, rhrx, hclx, hcrx, hfx, prm.moncen);
        dMy[v] = fb_slope<B>(Wm[v], W0[v], Wp[v], rhly, rhry, hcly, hcry, hfy, prm.moncen);
        Sx[v] = fb_half<B>(dMx[v], hfx);
        Sy[v] = fb_half<B>(dMy[v], hfy);
      }
      fb_hancock<B>(W0, dMx, dMy, gamma, dtib, Mp);
      // lane + 1's lower x-face state, its theta
      B WRx[4];
  #pragma unroll
      for (int v = 0; v < 4; ++v) WRx[v] = from_right(Mp[v] - Sx[v]);
      const double thr = from_right(thk);

      if (k >= y0) {
        // y-face k, between rows k - 1 and k
        double tf = godunov ? 1.0 : nmax(thm, thk);
        if (fyin && !(tf == 0.0)) {
          B WL[4], WR[4], UL[4], UR[4], Ff[4];
  #pragma unroll
          for (int v = 0; v < 4; ++v) {
            WL[v] = Qm[v];
            WR[v] = Mp[v] - Sy[v];
          }
          cons<2>(WL, UL, 1, gamma);
          cons<2>(WR, UR, 1, gamma);
          riemann<2, RS>(WL, WR, UL, UR, Ff, 1, gamma, mc2, prm.lowmach_fv);
          const B tb = (B)tf;
  #pragma unroll
          for (int v = 0; v < 4; ++v) Fyk[v] = fb_blend<B>(tb, Ff[v], Fyk[v]);
        }
        // update of row k - 1
        B FxL[4];
  #pragma unroll
        for (int v = 0; v < 4; ++v) FxL[v] = from_left(FxR[v]);
        const double tfxL = from_left(tfxR);
        if (k - 1 >= y0 && out) {
          const bool touched =
              godunov || !(tfxL == 0.0) || !(tfxR == 0.0) || !(tfyB == 0.0) || !(tf == 0.0);
          if (touched) {
            const int y = k - 1;
            const double rdy = g.rhfp[1][y + NGH];
            const long long ci = (long long)y * Nx + x;
  #pragma unroll
            for (int v = 0; v < 4; ++v)
              Uout[v * g.Nc + ci] = fb_apply(Ucv[v * g.Nc + ci], (double)FxL[v], (double)FxR[v],
                                             (double)FyB[v], (double)Fyk[v], rdx, rdy, dti);
          }
        }
  #pragma unroll
        for (int v = 0; v < 4; ++v) FyB[v] = Fyk[v];
        tfyB = tf;
      }
      if (k >= y0 && k < y1) {
        // x-face (k, x + 1/2)
        double tf = godunov ? 1.0 : nmax(thk, thr);
        if (fxin && !(tf == 0.0)) {
          B WL[4], UL[4], UR[4], Ff[4];
  #pragma unroll
          for (int v = 0; v < 4; ++v) WL[v] = Mp[v] + Sx[v];
          cons<2>(WL, UL, 0, gamma);
          cons<2>(WRx, UR, 0, gamma);
          riemann<2, RS>(WL, WRx, UL, UR, Ff, 0, gamma, mc2, prm.lowmach_fv);
          const B tb = (B)tf;
  #pragma unroll
          for (int v = 0; v < 4; ++v) Fxk[v] = fb_blend<B>(tb, Ff[v], Fxk[v]);
        }
  #pragma unroll
        for (int v = 0; v < 4; ++v) FxR[v] = Fxk[v];
        tfxR = tf;
      }
  #pragma unroll
      for (int v = 0; v < 4; ++v) {
        Qm[v] = Mp[v] + Sy[v];
        Wm[v] = W0[v];
        W0[v] = Wp[v];
      }
      thm = thk;
    }
  };
  // bit i of wnd: a busy row among bits i .. i + 4 of busy (row k: i = k - y0 + 1)
  const unsigned long long wnd = busy | (busy >> 1) | (busy >> 2) | (busy >> 3) | (busy >> 4);
  const int nk = y1 - y0 + 2;  // rows k = y0 - 1 .. y1
  const unsigned long long all = (nk >= 64) ? ~0ull : ((1ull << nk) - 1);
  if ((wnd & all) == all)
    march(std::false_type{});
  else
    march(std::true_type{});
}
template <int RS, typename B, bool STORED>
__global__ void __launch_bounds__(MW * MWAVES)
k_fbmarch2d(Geom g, Params prm, const double* __restrict__ Ucv,
            const double* __restrict__ theta, const B* __restrict__ Fx,
            const B* __restrict__ Fy, const double* __restrict__ dt_dev, double w_i,
            int MR, double* __restrict__ Uout, const int* __restrict__ gate) {
  k_fbmarch2d_body<RS, B, STORED>(hw_block(), g, prm, Ucv, theta, Fx, Fy, dt_dev, w_i, MR, Uout, gate);
}

// ---------------------------------------------------------------------------
// High-order fluxes of time node t on the FV faces of both dimensions
// (src/sdader_simulator.py:288-323, src/sd_simulator.py:208-212).  A block
// owns EPB consecutive elements: their predictor slab of node t is staged in
// LDS with coalesced loads, then one work item per (dimension, element, flux
// point) evaluates the flux on the element's lines, averages it over the
// transverse subcells and writes the face values.
// ---------------------------------------------------------------------------
// Element stride of the LDS slab: NL + 1 values, odd, so that the lanes of a
// wave, which read up to 13 different elements at the same point, fall into
// different LDS banks (NL is a multiple of 16 at p = 3).
__host__ __device__ constexpr int ho_nlp(int P) { return (P + 1) * (P + 1) + 1; }

template <int P, int A, typename B>
__device__ __forceinline__ void ho2d_items(const Ops<P, B>& op, const Geom& g, const Params& prm,
                                           int t, const B* __restrict__ Ff, B* __restrict__ Fh,
                                           const B* slab, long long e0, int ne, int EPB);

template <int P, typename B>
__global__ void __launch_bounds__(256)
k_ho2d(Ops<P, B> op, Geom g, Params prm, int t, const B* __restrict__ Uad,
       const B* __restrict__ Ffx, const B* __restrict__ Ffy, B* __restrict__ Fhx,
       B* __restrict__ Fhy, int EPB) {
  SPD_SKIP(g);
  constexpr int V = 4, n = P + 1, nf = P + 2, NL = n * n;
  B* slab = reinterpret_cast<B*>(spd_dyn_lds);  // [V][EPB][ho_nlp(P)]
  const long long e0 = (long long)blockIdx.x * EPB;
  const int ne = (int)((g.NE - e0 < EPB) ? (g.NE - e0) : EPB);
  for (int k = threadIdx.x; k < V * ne * NL; k += blockDim.x) {
    const int v = k / (ne * NL), r = k % (ne * NL);
    slab[(v * EPB + r / NL) * ho_nlp(P) + r % NL] =
        Uad[((long long)(t * V + v) * g.NE + e0) * NL + r];
  }
  __syncthreads();
  ho2d_items<P, 0, B>(op, g, prm, t, Ffx, Fhx, slab, e0, ne, EPB);
  ho2d_items<P, 1, B>(op, g, prm, t, Ffy, Fhy, slab, e0, ne, EPB);
}

template <int P, int A, typename B>
__device__ __forceinline__ void ho2d_items(const Ops<P, B>& op, const Geom& g, const Params& prm,
                                           int t, const B* __restrict__ Ff, B* __restrict__ Fh,
                                           const B* slab, long long e0, int ne, int EPB) {
  constexpr int V = 4, n = P + 1, nf = P + 2, NL = n * n;
  constexpr int a = A;
  const int Nx = g.N[0], Ny = g.N[1];
  for (int item = threadIdx.x; item < ne * nf; item += blockDim.x) {
    // x: consecutive items walk the flux points of an element (contiguous x
    // faces); y: the elements at one flux point (the lanes of a wave share
    // the face-or-interior branch and write along one row of y faces)
    const int el = (a == 0) ? item / nf : item % ne;
    const int f = (a == 0) ? item % nf : item / ne;
    const long long e = e0 + el;
    const int cx = (int)(e % Nx), cy = (int)(e / Nx);
    const int ca = (a == 0) ? cx : cy;
    const int Na = (a == 0) ? Nx : Ny;
    if (f == nf - 1 && ca != Na - 1) continue;
    B Fl[n][V];
    if (f == 0 || f == nf - 1) {
      // element-face Riemann flux on the face grid of dimension a
      const int fa = ca + (f == 0 ? 0 : 1);
      const long long plane = (a == 0) ? (long long)(Nx + 1) * Ny * n : (long long)Nx * n * (Ny + 1);
#pragma unroll
      for (int lp = 0; lp < n; ++lp) {
        const long long i0 = (a == 0) ? ((long long)cy * n + lp) * (Nx + 1) + fa
                                      : (long long)fa * Nx * n + (long long)cx * n + lp;
#pragma unroll
        for (int v = 0; v < V; ++v) Fl[lp][v] = Ff[(long long)(t * V + v) * plane + i0];
      }
    } else {
#pragma unroll
      for (int lp = 0; lp < n; ++lp) {
        B M[V], W[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const B* u = slab + (v * EPB + el) * ho_nlp(P);
          B acc = B(0);
#pragma unroll
          for (int s = 0; s < n; ++s) {
            const int l = (a == 0) ? (lp * n + s) : (s * n + lp);
            acc += op.s2f[f][s] * u[l];
          }
          M[v] = acc;
        }
        prims<2>(M, W, a, (B)prm.gamma);
        flux<2>(W, Fl[lp], a, (B)prm.gamma);
      }
    }
    const long long nxf = (a == 0) ? (long long)Nx * n + 1 : (long long)Nx * n;
    const long long nface = nxf * ((a == 0) ? (long long)Ny * n : (long long)Ny * n + 1);
#pragma unroll
    for (int q = 0; q < n; ++q) {
      B Fa[V];
#pragma unroll
      for (int v = 0; v < V; ++v) {
        B acc = B(0);
#pragma unroll
        for (int s = 0; s < n; ++s) acc += op.s2c[q][s] * Fl[s][v];
        Fa[v] = acc;
      }
      const long long fx = (a == 0) ? (long long)cx * n + f : (long long)cx * n + q;
      const long long fy = (a == 0) ? (long long)cy * n + q : (long long)cy * n + f;
#pragma unroll
      for (int v = 0; v < V; ++v) Fh[v * nface + fy * nxf + fx] = Fa[v];
    }
  }
}


// ---------------------------------------------------------------------------
// Flux at the interior flux point f (1..p) of line lp of one element along
// dimension A: Lagrange interpolation of the line's n solution points, then
// primitives and the Euler flux (src/sd_simulator.py:190-198,
// src/hydro.py:127-170); the arithmetic of ho2d_items.  ld(v, l) is the
// element's value of variable v at solution point l (x fastest).
// ---------------------------------------------------------------------------
template <int P, int A, typename B, typename Ld>
__device__ __forceinline__ void ho_fp_flux(const Ops<P, B>& op, B gamma, int f, int lp, Ld&& ld,
                                           B* Fl) {
  constexpr int V = 4, n = P + 1;
  B M[V], W[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    B acc = B(0);
#pragma unroll
    for (int s = 0; s < n; ++s) {
      const int l = (A == 0) ? (lp * n + s) : (s * n + lp);
      acc = fma(op.s2f[f][s], ld(v, l), acc);
    }
    M[v] = acc;
  }
  prims_rn<2>(M, W, A, gamma);
  flux_rn<2>(W, Fl, A, gamma);
}

// ---------------------------------------------------------------------------
// High-order FV face fluxes of time node t and the candidate state in one
// pass (src/sdader_simulator.py:288-323,335-348, src/fv_simulator.py:189-205).
// A block of 256 threads owns HocTile<P>::E elements (about 256 cells): the
// state of its cells and the element-face Riemann fluxes it needs are loaded
// into registers first, then the predictor slab of node t is staged in LDS;
// one work item per (dimension, element, face column, line) evaluates the
// flux on the line (element faces: the Riemann flux of the face grid), one
// per (dimension, element, face column, transverse subcell) averages the
// lines, and every cell of the block's elements gets
// U_out = U_cv - dt_i sum(dF/dx), the expression of k_apply on the same face
// values.  The FV face fluxes stay in LDS: each element forms its own n + 1
// face columns per dimension (a face shared by two elements is formed by
// both, identically).  U_out may be U_cv (the update in place without the
// fallback): a cell is read and written by the same thread only.  A block
// owns a run of at most E elements of one element row.
// With DET (and flags != null) the kernel also runs the first half of the
// trouble detection (src/trouble_detection.py:5-55) on its cells: the NAD
// bounds from the old primitives of the cells and their one-cell ring (staged
// in LDS before the slab), the NAD and PAD tests of the candidate, the global
// NAD gate (any_flag, when not null), and per cell the final flag, or the code
// that k_sed2d finishes when the smooth-extrema relaxation follows.
//   LDS: slab [V][E][ho_nlp], face columns [2][E][nf][n][V] (lines, then
//   their transverse averages).
// ---------------------------------------------------------------------------
template <int P> struct HocTile {
  static constexpr int n = P + 1, nf = P + 2, NL = n * n;
  static constexpr int E = (256 / NL > 0) ? 256 / NL : 1;
  static constexpr int NC = E * NL;               // cells of a block
  static constexpr int NI = E * nf * n;           // face-column items per dimension
  static constexpr int CPT = (NC + 255) / 256;    // cells per thread
  static constexpr int IPT = (2 * NI + 255) / 256;  // items per thread (both dimensions)
  template <typename B> static constexpr int lds_bytes() {
    return (4 * E * ho_nlp(P) + 2 * NI * 4) * (int)sizeof(B);
  }
  // with the detection's first half: the old primitives of the block's cells
  // and their ring, (nvar, n + 2, E n + 2) doubles (nvar = 2, or 4 with the
  // global gate), go to the slab's LDS (free once the line fluxes are formed)
  // when they fit there, else after the face columns
  template <typename B> static constexpr int ring_bytes(int nvar) {
    return nvar * (n + 2) * (E * n + 2) * 8;
  }
  template <typename B> static constexpr bool ring_in_slab(int nvar) {
    return ring_bytes<B>(nvar) <= 4 * E * ho_nlp(P) * (int)sizeof(B);
  }
  template <typename B> static constexpr int lds_bytes_det(int nvar) {
    return ring_in_slab<B>(nvar) ? lds_bytes<B>() : lds_bytes<B>() + ring_bytes<B>(nvar);
  }
};

// FULL: every block holds E elements (E per block and a row length that is a
// multiple of E); the index arithmetic then divides by constants only.
template <int P, typename B, bool DET, bool FULL = false>
__device__ __forceinline__ void k_hocand2d_body(const BlockIx bi, Ops<P, B> op, Geom g, Params prm, int t, const B* __restrict__ Uad,
           const B* __restrict__ Ffx, const B* __restrict__ Ffy, const double* Ucv,
           const double* __restrict__ dt_dev, double w_i, double* Uout, B* __restrict__ Fhx,
           B* __restrict__ Fhy, uint8_t* __restrict__ flags, int* __restrict__ any_flag,
           const int* __restrict__ fh_gate, int Er) {
  SPD_SKIP(g);
  // the face-flux arrays are written when given and, with a gate, when it is set
  if (fh_gate && *fh_gate == 0) Fhx = Fhy = nullptr;
  using HT = HocTile<P>;
  constexpr int V = 4, n = P + 1, nf = P + 2, NL = n * n, E = HT::E, NI = HT::NI;
  B* slab = reinterpret_cast<B*>(spd_dyn_lds);  // [V][E][ho_nlp]
  B* col = slab + V * E * ho_nlp(P);            // [2][E][nf][n][V]
  const int Nx = g.N[0], Ny = g.N[1], Nsx = g.Ns[0];
  // the block's elements: a run of at most Eb <= E elements of one element
  // row (Er: the elements per block chosen at run time, 0 = E)
  const int Eb = FULL ? E : ((Er > 0 && Er < E) ? Er : E);
  const int nseg = FULL ? Nx / E : (Nx + Eb - 1) / Eb;
  const int ey = bi.x / nseg, ex0 = (bi.x % nseg) * Eb;
  const long long e0 = (long long)ey * Nx + ex0;
  const int ne = FULL ? E : ((Nx - ex0 < Eb) ? (Nx - ex0) : Eb);
  const int tid = threadIdx.x;
  const B gamma = (B)prm.gamma;
  const double dti = (*dt_dev) * w_i;
  // the block's cells row by row across its elements (consecutive threads
  // take consecutive x when the elements are consecutive in a row)
  auto cell_of = [&](int item, int& el, int& i, int& qy, long long& ci, int& x, int& y) {
    qy = item / (ne * n);
    const int xl = item % (ne * n);
    el = xl / n;
    i = xl % n;
    x = (ex0 + el) * n + i;
    y = ey * n + qy;
    ci = (long long)y * Nsx + x;
  };
  // 1. loads into registers: the cells' state, the element-face fluxes
  double u[HT::CPT][V];
#pragma unroll
  for (int c = 0; c < HT::CPT; ++c) {
    const int item = tid + c * 256;
    if (item < ne * NL) {
      int el, i, qy, x, y;
      long long ci;
      cell_of(item, el, i, qy, ci, x, y);
#pragma unroll
      for (int v = 0; v < V; ++v) u[c][v] = Ucv[v * g.Nc + ci];
    }
  }
  B fl[HT::IPT][V];
#pragma unroll
  for (int it = 0; it < HT::IPT; ++it) {
    const int item = tid + it * 256;
    const int a = item / NI, r = item % NI;
    const int lp = r % n, f = (r / n) % nf, el = r / (n * nf);
    if (item >= 2 * NI || el >= ne || (f != 0 && f != nf - 1)) continue;
    const int cx = ex0 + el, cy = ey;
    if (a == 0) {
      const long long i0 = ((long long)cy * n + lp) * (Nx + 1) + cx + (f == 0 ? 0 : 1);
      const long long plane = (long long)(Nx + 1) * Ny * n;
#pragma unroll
      for (int v = 0; v < V; ++v) fl[it][v] = Ffx[(long long)(t * V + v) * plane + i0];
    } else {
      const long long i0 = (long long)(cy + (f == 0 ? 0 : 1)) * Nx * n + (long long)cx * n + lp;
      const long long plane = (long long)Nx * n * (Ny + 1);
#pragma unroll
      for (int v = 0; v < V; ++v) fl[it][v] = Ffy[(long long)(t * V + v) * plane + i0];
    }
  }
  const bool det = DET && flags != nullptr;
  const int nvar_nad = any_flag ? 4 : 2;  // velocities only feed the global gate
  // 2. the predictor slab of node t
  for (int k = tid; k < V * ne * NL; k += 256) {
    const int v = k / (ne * NL), r = k % (ne * NL);
    slab[(v * E + r / NL) * ho_nlp(P) + r % NL] = Uad[((long long)(t * V + v) * g.NE + e0) * NL + r];
  }
  __syncthreads();
  // 3. line fluxes of every face column, both dimensions
#pragma unroll
  for (int it = 0; it < HT::IPT; ++it) {
    const int item = tid + it * 256;
    const int a = item / NI, r = item % NI;
    const int lp = r % n, f = (r / n) % nf, el = r / (n * nf);
    if (item >= 2 * NI || el >= ne) continue;
    B F[V];
    if (f == 0 || f == nf - 1) {
#pragma unroll
      for (int v = 0; v < V; ++v) F[v] = fl[it][v];
    } else {
      auto ld = [&](int v, int l) { return slab[(v * E + el) * ho_nlp(P) + l]; };
      if (a == 0)
        ho_fp_flux<P, 0, B>(op, gamma, f, lp, ld, F);
      else
        ho_fp_flux<P, 1, B>(op, gamma, f, lp, ld, F);
    }
#pragma unroll
    for (int v = 0; v < V; ++v) col[item * V + v] = F[v];
  }
  __syncthreads();
  // 4. transverse averages (through registers: they replace the lines)
  B av[HT::IPT][V];
#pragma unroll
  for (int it = 0; it < HT::IPT; ++it) {
    const int item = tid + it * 256;
    if (item >= 2 * NI || (item % NI) / (n * nf) >= ne) continue;
    const int q = item % n, base = item - q;  // lines of the same (a, el, f)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      B acc = B(0);
#pragma unroll
      for (int s2 = 0; s2 < n; ++s2) acc = fma(op.s2c[q][s2], col[(base + s2) * V + v], acc);
      av[it][v] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < HT::IPT; ++it) {
    const int item = tid + it * 256;
    if (item >= 2 * NI || (item % NI) / (n * nf) >= ne) continue;
#pragma unroll
    for (int v = 0; v < V; ++v) col[item * V + v] = av[it][v];
    if (Fhx) {
      // the FV face-flux arrays for a fallback that reads them: an element
      // writes its first n face columns, the last element of a row also its
      // last (the layout of k_ho2d)
      const int a = item / NI, r = item % NI;
      const int q = r % n, f = (r / n) % nf, el = r / (n * nf);
      const int cx = ex0 + el, cy = ey;
      const int ca = (a == 0) ? cx : cy, Na = (a == 0) ? Nx : Ny;
      if (f == nf - 1 && ca != Na - 1) continue;
      const long long nxf = (a == 0) ? (long long)Nx * n + 1 : (long long)Nx * n;
      const long long nface = nxf * ((a == 0) ? (long long)Ny * n : (long long)Ny * n + 1);
      const long long fx = (a == 0) ? (long long)cx * n + f : (long long)cx * n + q;
      const long long fy = (a == 0) ? (long long)cy * n + q : (long long)cy * n + f;
      B* Fh = (a == 0) ? Fhx : Fhy;
#pragma unroll
      for (int v = 0; v < V; ++v) Fh[v * nface + fy * nxf + fx] = av[it][v];
    }
  }
  __syncthreads();
  // 4b. (DET) the old primitives of the block's cells and their one-cell ring
  // (density and pressure, all variables with the gate; the staging of
  // k_detect2d), in the slab's LDS when it is large enough, else after the
  // face columns
  const int RX = ne * n + 2, RN = RX * (n + 2);
  double* wo = reinterpret_cast<double*>(
      spd_dyn_lds + (HT::template ring_in_slab<B>(nvar_nad) ? 0 : HT::template lds_bytes<B>()));
  if (det) {
    // the block's own cells from the registers, the ring from memory
#pragma unroll
    for (int c = 0; c < HT::CPT; ++c) {
      const int item = tid + c * 256;
      if (item >= ne * NL) continue;
      const int o = (item / (ne * n) + 1) * RX + item % (ne * n) + 1;
      double W[4];
      prims<2>(u[c], W, 0, prm.gamma);
      wo[o] = W[0];
      wo[RN + o] = W[3];
      if (nvar_nad == 4) {
        wo[2 * RN + o] = W[1];
        wo[3 * RN + o] = W[2];
      }
    }
    const int X0 = ex0 * n - 1, Y0 = ey * n - 1;
    const int RY = n + 2, NB = 2 * RX + 2 * (RY - 2);
    for (int j = tid; j < NB; j += 256) {
      int lx, ly;
      if (j < 2 * RX) {
        lx = j % RX;
        ly = (j < RX) ? 0 : RY - 1;
      } else {
        lx = ((j - 2 * RX) % 2) * (RX - 1);
        ly = 1 + (j - 2 * RX) / 2;
      }
      const int k = ly * RX + lx;
      const int x = X0 + lx, y = Y0 + ly;
      double uu[4], W[4] = {qnan(), qnan(), qnan(), qnan()};
      bool fl2[2];
      if (stage_cons(Ucv, g, x, y, uu, fl2)) {
        prims<2>(uu, W, 0, prm.gamma);
        if (fl2[0]) W[1] = -W[1];
        if (fl2[1]) W[2] = -W[2];
        int c3[3] = {x, y, 0};
        stored_ghost<2>(g, c3, W);  // ghosts of "ic" / "pressure" sides
      }
      wo[k] = W[0];
      wo[RN + k] = W[3];
      if (nvar_nad == 4) {
        wo[2 * RN + k] = W[1];
        wo[3 * RN + k] = W[2];
      }
    }
    __syncthreads();
  }
  // 5. the candidate; (DET) its NAD / PAD test: the code PAD (bit 0) | density
  // candidate (bit 1) | pressure candidate (bit 2) for k_sed2d when the
  // smooth-extrema relaxation follows, else the final flag
  const bool sed = prm.p > 1 && prm.sed;
  bool cand = false;
#pragma unroll
  for (int c = 0; c < HT::CPT; ++c) {
    const int item = tid + c * 256;
    if (item >= ne * NL) continue;
    int el, i, qy, x, y;
    long long ci;
    cell_of(item, el, i, qy, ci, x, y);
    const double rdx = g.rhfp[0][x + NGH], rdy = g.rhfp[1][y + NGH];
    const B* fx0 = col + ((el * nf + i) * n + qy) * V;         // x-face i, subcell row qy
    const B* fx1 = fx0 + n * V;                                // x-face i + 1
    const B* fy0 = col + (NI + (el * nf + qy) * n + i) * V;    // y-face qy, subcell column i
    const B* fy1 = fy0 + n * V;                                // y-face qy + 1
    double un[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      double d = u[c][v] * 0.0;
      d = d + ((double)fx1[v] - (double)fx0[v]) * rdx;
      d = d + ((double)fy1[v] - (double)fy0[v]) * rdy;
      un[v] = u[c][v] - d * dti;
      Uout[v * g.Nc + ci] = un[v];
    }
    if (det) {
      double Wn[4];
      prims<2>(un, Wn, 0, prm.gamma);
      bool pt[4] = {false, false, false, false};
      const int o = (qy + 1) * RX + el * n + i + 1;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k >= nvar_nad) break;
        const int v = (k == 0) ? 0 : (k == 1 ? 3 : k - 1);
        const double* w = wo + k * RN + o;
        double mx = w[0], mn = mx;
        const double nb[4] = {w[-1], w[1], w[-RX], w[RX]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          mx = nmax(mx, nb[q]);
          mn = nmin(mn, nb[q]);
        }
        double lo = mn, hi = mx;
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
      bool pad = false;
      if (prm.pad)
        pad = !(Wn[0] >= prm.min_rho) || !(Wn[0] <= prm.max_rho) || !(Wn[3] >= prm.min_P);
      flags[ci] = sed ? (uint8_t)((pad ? 1 : 0) | (pt[0] ? 2 : 0) | (pt[3] ? 4 : 0))
                      : (uint8_t)((pad || pt[0] || pt[3]) ? 1 : 0);
    }
  }
  if (det && any_flag) {
    const int blk = __syncthreads_or(cand ? 1 : 0);
    if (blk && threadIdx.x == 0 && *(volatile int*)any_flag == 0) atomicOr(any_flag, 1);
  }
}
template <int P, typename B, bool DET, bool FULL>
__global__ void __launch_bounds__(256)
k_hocand2d(Ops<P, B> op, Geom g, Params prm, int t, const B* __restrict__ Uad,
           const B* __restrict__ Ffx, const B* __restrict__ Ffy, const double* Ucv,
           const double* __restrict__ dt_dev, double w_i, double* Uout, B* __restrict__ Fhx,
           B* __restrict__ Fhy, uint8_t* __restrict__ flags, int* __restrict__ any_flag,
           const int* __restrict__ fh_gate, int Er) {
  k_hocand2d_body<P, B, DET, FULL>(hw_block(), op, g, prm, t, Uad, Ffx, Ffy, Ucv, dt_dev, w_i, Uout, Fhx, Fhy, flags, any_flag, fh_gate, Er);
}

// ---------------------------------------------------------------------------
// Compacted fallback.  Element-aligned cell tiles of FbTile<P> (EX x EY
// elements, at most 32 x 8 cells): k_fbtile_list appends to `list` every tile
// one of whose cells borders a face with a non-zero blending coefficient
// (theta of the tile and of its one-cell ring without the corners; every
// tile with godunov), and k_fbtile2d runs the MUSCL-Hancock + Riemann
// fallback on the listed tiles only: as k_fbcorrect2d, with the high-order
// face fluxes formed in LDS from the predictor slab of node t and the
// element-face Riemann fluxes (the arithmetic of k_hocand2d) instead of read
// from HBM, and only the cells that border a blended face written (the other
// cells keep the candidate).  counters[0] counts the list, counters[1] hands
// out its entries to the blocks (both zeroed before k_fbtile_list).
// ---------------------------------------------------------------------------
template <int P> struct FbTile {
  static constexpr int n = P + 1;
  static constexpr int EX = (32 / n > 0) ? 32 / n : 1;
  static constexpr int EY = (8 / n > 0) ? 8 / n : 1;
  static constexpr int TXc = EX * n, TYc = EY * n;
  static constexpr int CXc = TXc + 2, CYc = TYc + 2;   // tile + 1 halo
  static constexpr int WXc = TXc + 4, WYc = TYc + 4;   // tile + 2 halo
  static constexpr int NFX = TYc * (TXc + 1), NFY = (TYc + 1) * TXc;  // faces of the tile
};
template <int P> inline long long fbtile_count(const Geom& g) {
  using FT = FbTile<P>;
  return (long long)((g.N[0] + FT::EX - 1) / FT::EX) * ((g.N[1] + FT::EY - 1) / FT::EY);
}

__global__ void __launch_bounds__(256) k_fill_u8(uint8_t* p, long long n, uint8_t v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_zero2(int* c) {
  if (threadIdx.x < 2 && blockIdx.x == 0) c[threadIdx.x] = 0;
}

// Path of the adaptive form for one substep, before its high-order flux
// kernel: counters[2] = 1 (column strips over the face-flux arrays) when
// `force` is set or the previous substep listed more than `thr` active
// tiles, else 0 (compacted tiles, no face-flux arrays); the list counters are
// cleared.  A replayed tape (no tile marks) forces the column strips.
__global__ void k_fb_decide(Geom g, int* counters, long long thr, int force) {
  SPD_SKIP(g);
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    counters[2] = (force || counters[0] > thr) ? 1 : 0;
    counters[0] = 0;
    counters[1] = 0;
  }
}

// Active-tile list from the marks k_theta2d set (every tile with godunov):
// one thread per tile, which clears its mark; blocks of 256 threads.
__global__ void __launch_bounds__(256)
k_fbmark_list(Geom g, Params prm, uint8_t* __restrict__ marks, int* __restrict__ counters,
              int* __restrict__ list, long long ntiles) {
  SPD_SKIP(g);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool live = false;
  if (t < ntiles) {
    const bool m = marks[t] != 0;
    if (m) marks[t] = 0;
    live = m || prm.godunov != 0;
  }
  // one atomic per block reserves its entries (same-address atomics
  // serialize in L2); the waves take consecutive ranges
  __shared__ int s_cnt[4];
  __shared__ int s_base;
  const unsigned long long b = __ballot(live);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) s_cnt[wave] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    s_base = tot ? atomicAdd(&counters[0], tot) : 0;
  }
  __syncthreads();
  int base = s_base;
  for (int w = 0; w < wave; ++w) base += s_cnt[w];
  if (live) list[base + __popcll(b & ((1ull << lane) - 1ull))] = (int)t;
}

// One wave per tile, FBL_WAVES tiles (consecutive in x) per block; one
// atomic per block reserves the block's entries (same-address atomics
// serialize in L2).
constexpr int FBL_WAVES = 8;

template <int P>
__global__ void __launch_bounds__(64 * FBL_WAVES)
k_fbtile_list(Geom g, Params prm, const double* __restrict__ theta, int* __restrict__ counters,
              int* __restrict__ list, long long ntiles) {
  SPD_SKIP(g);
  using FT = FbTile<P>;
  __shared__ int s_live[FBL_WAVES];
  __shared__ int s_base;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int ntx = (g.N[0] + FT::EX - 1) / FT::EX;
  const long long tile = (long long)blockIdx.x * FBL_WAVES + wave;
  const int Nx = g.Ns[0], Ny = g.Ns[1], NGx = g.Ng[0];
  bool live = false;
  if (tile < ntiles) {
    live = prm.godunov != 0;
    const int x0 = (int)(tile % ntx) * FT::TXc, y0 = (int)(tile / ntx) * FT::TYc;
    constexpr int NIT = (FT::CXc * FT::CYc + 63) / 64;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int k = lane + it * 64;
      const int lx = k % FT::CXc, ly = k / FT::CXc;
      const int x = x0 + lx - 1, y = y0 + ly - 1;
      const bool corner = (lx == 0 || lx == FT::CXc - 1) && (ly == 0 || ly == FT::CYc - 1);
      if (k < FT::CXc * FT::CYc && !corner && x >= -1 && x <= Nx && y >= -1 && y <= Ny)
        live |= !(theta[(long long)(y + NGH) * NGx + (x + NGH)] == 0.0);
    }
  }
  const int any = __any(live) ? 1 : 0;
  if (lane == 0) s_live[wave] = any;
  __syncthreads();
  if (threadIdx.x == 0) {
    int cnt = 0;
    for (int w = 0; w < FBL_WAVES; ++w) cnt += s_live[w];
    s_base = cnt ? atomicAdd(&counters[0], cnt) : 0;
  }
  __syncthreads();
  if (lane == 0 && any) {
    int off = 0;
    for (int w = 0; w < wave; ++w) off += s_live[w];
    list[s_base + off] = (int)tile;
  }
}

template <int P, int RS, typename B>
__global__ void __launch_bounds__(256)
k_fbtile2d(Ops<P, B> op, Geom g, Params prm, int t, const B* __restrict__ Uad,
           const B* __restrict__ Ffx, const B* __restrict__ Ffy, const double* __restrict__ Ucv,
           const double* __restrict__ theta, int* __restrict__ counters,
           const int* __restrict__ list, const double* __restrict__ dt_dev, double w_i,
           double* __restrict__ Uout, bool adaptive) {
  SPD_SKIP(g);
  if (adaptive && counters[2] != 0) return;  // the adaptive form took the column strips
  using FT = FbTile<P>;
  constexpr int V = 4, n = P + 1, NL = n * n;
  constexpr int TXc = FT::TXc, TYc = FT::TYc, CXc = FT::CXc, CYc = FT::CYc;
  constexpr int WXc = FT::WXc, WYc = FT::WYc;
  constexpr int NFX = FT::NFX, NFY = FT::NFY;
  constexpr int SW = (4 * WYc * WXc > 4 * (NFX + NFY)) ? 4 * WYc * WXc : 4 * (NFX + NFY);
  static_assert(NFY >= NFX, "the x-face lines are staged in the y-face buffer");
  __shared__ B sW[SW];                 // W (4 x WY x WX), then the face fluxes
  __shared__ B sH[12][CYc][CXc];       // Mp(4), Sx(4), Sy(4)
  __shared__ double sT[CYc][CXc];
  __shared__ int s_idx;
  B* bFx = sW;                          // [TYc][TXc+1][4]
  B* bFy = sW + 4 * NFX;                // [TYc+1][TXc][4]
  const int Nx = g.Ns[0], Ny = g.Ns[1], NGx = g.Ng[0];
  const int NEx = g.N[0], NEy = g.N[1];
  const int ntx = (NEx + FT::EX - 1) / FT::EX;
  const B gamma = (B)prm.gamma;
  const double dti = (*dt_dev) * w_i;
  const B dtib = (B)dti;
  const bool godunov = prm.godunov != 0;
  const int count = counters[0];
  B(*W)[WYc][WXc] = reinterpret_cast<B(*)[WYc][WXc]>(sW);
  const long long planex = (long long)(NEx + 1) * NEy * n;
  const long long planey = (long long)NEx * n * (NEy + 1);
  auto U_at = [&](long long e, int v, int l) -> B {
    return Uad[((long long)(t * V + v) * g.NE + e) * NL + l];
  };

  for (;;) {
    __syncthreads();  // the previous tile's LDS is no longer read
    if (threadIdx.x == 0) s_idx = atomicAdd(&counters[1], 1);
    __syncthreads();
    const int li = s_idx;
    if (li >= count) break;
    const int tile = list[li];
    const int ex0 = (tile % ntx) * FT::EX, ey0 = (tile / ntx) * FT::EY;
    const int x0 = ex0 * n, y0 = ey0 * n;

    // 0. theta on tile + 1 halo
    for (int k = threadIdx.x; k < CXc * CYc; k += blockDim.x) {
      const int lx = k % CXc, ly = k / CXc;
      const int x = x0 + lx - 1, y = y0 + ly - 1;
      double th = 0.0;
      if (x >= -1 && x <= Nx && y >= -1 && y <= Ny)
        th = theta[(long long)(y + NGH) * NGx + (x + NGH)];
      sT[ly][lx] = th;
    }
    // 1. old primitives on tile + 2 halo
    for (int k = threadIdx.x; k < WXc * WYc; k += blockDim.x) {
      const int lx = k % WXc, ly = k / WXc;
      const int x = x0 + lx - 2, y = y0 + ly - 2;
      double Wv[4] = {qnan(), qnan(), qnan(), qnan()};
      if (in_ghost_range(g, x, y)) {
        int c[3] = {x, y, 0};
        load_W<2>(Ucv, g, c, Wv, prm.gamma);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) W[v][ly][lx] = (B)Wv[v];
    }
    __syncthreads();

    // 2. Hancock state and half slopes of the cells that border an active face
    for (int k = threadIdx.x; k < CXc * CYc; k += blockDim.x) {
      const int lx = k % CXc, ly = k / CXc;
      const int x = x0 + lx - 1, y = y0 + ly - 1;
      bool need = false;
      if (ly >= 1 && ly <= TYc && y < Ny) {
        if (lx <= TXc && x + 1 <= Nx) need |= godunov || !(nmax(sT[ly][lx], sT[ly][lx + 1]) == 0.0);
        if (lx >= 1 && x <= Nx) need |= godunov || !(nmax(sT[ly][lx - 1], sT[ly][lx]) == 0.0);
      }
      if (lx >= 1 && lx <= TXc && x < Nx) {
        if (ly <= TYc && y + 1 <= Ny) need |= godunov || !(nmax(sT[ly][lx], sT[ly + 1][lx]) == 0.0);
        if (ly >= 1 && y <= Ny) need |= godunov || !(nmax(sT[ly - 1][lx], sT[ly][lx]) == 0.0);
      }
      if (!need) continue;
      const int wx = lx + 1, wy = ly + 1;
      B Wc[4], dMh[2][4], Mp[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) Wc[v] = W[v][wy][wx];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int cb = (b == 0) ? x : y;
        const int G = cb + NGH;
        const B rhl = (B)g.rhcv[b][G - 1], rhr = (B)g.rhcv[b][G];
        const B hf = (B)g.hfp[b][G];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const B Wm = (b == 0) ? W[v][wy][wx - 1] : W[v][wy - 1][wx];
          const B Wp = (b == 0) ? W[v][wy][wx + 1] : W[v][wy + 1][wx];
          dMh[b][v] = fb_slope<B>(Wm, Wc[v], Wp, rhl, rhr, (B)g.hcv[b][G - 1], (B)g.hcv[b][G],
                                  hf, prm.moncen);
          sH[4 + 4 * b + v][ly][lx] = fb_half<B>(dMh[b][v], hf);
        }
      }
      fb_hancock<B>(Wc, dMh[0], dMh[1], gamma, dtib, Mp);
#pragma unroll
      for (int v = 0; v < 4; ++v) sH[v][ly][lx] = Mp[v];
    }
    __syncthreads();

    // 3. high-order face fluxes in the (now free) W buffer.  x: the line
    // fluxes of every face column (staged in the y-face part), then the
    // transverse averages into bFx.
    for (int k = threadIdx.x; k < NFX; k += blockDim.x) {
      const int fx = k % (TXc + 1), line = k / (TXc + 1);  // line = element row * n + lp
      const int x = x0 + fx, el_x = fx / n, f = fx % n;
      const int ey = ey0 + line / n, lp = line % n;
      if (x > Nx || ey >= NEy) continue;
      B Fl[V];
      if (f == 0) {  // element face fa = ex0 + el_x
        const long long i0 = ((long long)ey * n + lp) * (NEx + 1) + ex0 + el_x;
#pragma unroll
        for (int v = 0; v < V; ++v) Fl[v] = Ffx[(long long)(t * V + v) * planex + i0];
      } else {
        const long long e = (long long)ey * NEx + ex0 + el_x;
        ho_fp_flux<P, 0, B>(op, gamma, f, lp, [&](int v, int l) { return U_at(e, v, l); }, Fl);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) bFy[k * 4 + v] = Fl[v];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NFX; k += blockDim.x) {
      const int fx = k % (TXc + 1), ly = k / (TXc + 1);
      const int q = ly % n, row0 = ly - q;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        B acc = B(0);
#pragma unroll
        for (int s = 0; s < n; ++s)
          acc = fma(op.s2c[q][s], bFy[((row0 + s) * (TXc + 1) + fx) * 4 + v], acc);
        bFx[k * 4 + v] = acc;
      }
    }
    __syncthreads();
    // y: the line fluxes of every face row into bFy, the averages through
    // registers (they overwrite the lines)
    for (int k = threadIdx.x; k < NFY; k += blockDim.x) {
      const int col = k % TXc, fy = k / TXc;  // col = element column * n + lp
      const int y = y0 + fy, el_y = fy / n, f = fy % n;
      const int ex = ex0 + col / n, lp = col % n;
      if (y > Ny || ex >= NEx) continue;
      B Fl[V];
      if (f == 0) {  // element face fa = ey0 + el_y
        const long long i0 = (long long)(ey0 + el_y) * NEx * n + (long long)ex * n + lp;
#pragma unroll
        for (int v = 0; v < V; ++v) Fl[v] = Ffy[(long long)(t * V + v) * planey + i0];
      } else {
        const long long e = (long long)(ey0 + el_y) * NEx + ex;
        ho_fp_flux<P, 1, B>(op, gamma, f, lp, [&](int v, int l) { return U_at(e, v, l); }, Fl);
      }
#pragma unroll
      for (int v = 0; v < V; ++v) bFy[k * 4 + v] = Fl[v];
    }
    __syncthreads();
    constexpr int YIT = (NFY + 255) / 256;
    B avg[YIT][4];
#pragma unroll
    for (int it = 0; it < YIT; ++it) {
      const int k = threadIdx.x + it * 256;
      if (k >= NFY) continue;
      const int lx = k % TXc, fy = k / TXc;
      const int q = lx % n, col0 = lx - q;
#pragma unroll
      for (int v = 0; v < V; ++v) {
        B acc = B(0);
#pragma unroll
        for (int s = 0; s < n; ++s) acc = fma(op.s2c[q][s], bFy[(fy * TXc + col0 + s) * 4 + v], acc);
        avg[it][v] = acc;
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < YIT; ++it) {
      const int k = threadIdx.x + it * 256;
      if (k >= NFY) continue;
#pragma unroll
      for (int v = 0; v < V; ++v) bFy[k * 4 + v] = avg[it][v];
    }
    __syncthreads();

    // 4. fallback fluxes where the coefficient is non-zero, blended in place
    for (int k = threadIdx.x; k < NFX; k += blockDim.x) {
      const int fx = k % (TXc + 1), ly = k / (TXc + 1);
      const int x = x0 + fx, y = y0 + ly;
      if (x > Nx || y >= Ny) continue;
      const double th = godunov ? 1.0 : nmax(sT[ly + 1][fx], sT[ly + 1][fx + 1]);
      if (th == 0.0) continue;
      B WL[4], WR[4], UL[4], UR[4], Ff[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        WL[v] = sH[v][ly + 1][fx] + sH[4 + v][ly + 1][fx];
        WR[v] = sH[v][ly + 1][fx + 1] - sH[4 + v][ly + 1][fx + 1];
      }
      cons<2>(WL, UL, 0, gamma);
      cons<2>(WR, UR, 0, gamma);
      riemann<2, RS>(WL, WR, UL, UR, Ff, 0, gamma, (B)prm.min_c2, prm.lowmach_fv);
      const B tb = (B)th;
#pragma unroll
      for (int v = 0; v < 4; ++v) bFx[k * 4 + v] = fb_blend<B>(tb, Ff[v], bFx[k * 4 + v]);
    }
    for (int k = threadIdx.x; k < NFY; k += blockDim.x) {
      const int lx = k % TXc, fy = k / TXc;
      const int x = x0 + lx, y = y0 + fy;
      if (x >= Nx || y > Ny) continue;
      const double th = godunov ? 1.0 : nmax(sT[fy][lx + 1], sT[fy + 1][lx + 1]);
      if (th == 0.0) continue;
      B WL[4], WR[4], UL[4], UR[4], Ff[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        WL[v] = sH[v][fy][lx + 1] + sH[8 + v][fy][lx + 1];
        WR[v] = sH[v][fy + 1][lx + 1] - sH[8 + v][fy + 1][lx + 1];
      }
      cons<2>(WL, UL, 1, gamma);
      cons<2>(WR, UR, 1, gamma);
      riemann<2, RS>(WL, WR, UL, UR, Ff, 1, gamma, (B)prm.min_c2, prm.lowmach_fv);
      const B tb = (B)th;
#pragma unroll
      for (int v = 0; v < 4; ++v) bFy[k * 4 + v] = fb_blend<B>(tb, Ff[v], bFy[k * 4 + v]);
    }
    __syncthreads();

    // 5. the cells that border a blended face
    for (int k = threadIdx.x; k < TXc * TYc; k += blockDim.x) {
      const int tx = k % TXc, ty = k / TXc;
      const int x = x0 + tx, y = y0 + ty;
      if (x >= Nx || y >= Ny) continue;
      const bool touched = godunov || !(nmax(sT[ty + 1][tx], sT[ty + 1][tx + 1]) == 0.0) ||
                           !(nmax(sT[ty + 1][tx + 1], sT[ty + 1][tx + 2]) == 0.0) ||
                           !(nmax(sT[ty][tx + 1], sT[ty + 1][tx + 1]) == 0.0) ||
                           !(nmax(sT[ty + 1][tx + 1], sT[ty + 2][tx + 1]) == 0.0);
      if (!touched) continue;
      const double rdx = g.rhfp[0][x + NGH], rdy = g.rhfp[1][y + NGH];
      const long long ci = (long long)y * Nx + x;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        Uout[v * g.Nc + ci] = fb_apply(
            Ucv[v * g.Nc + ci], (double)bFx[(ty * (TXc + 1) + tx) * 4 + v],
            (double)bFx[(ty * (TXc + 1) + tx + 1) * 4 + v], (double)bFy[(ty * TXc + tx) * 4 + v],
            (double)bFy[((ty + 1) * TXc + tx) * 4 + v], rdx, rdy, dti);
    }
  }
}

}  // namespace fv2d
}  // namespace spd
