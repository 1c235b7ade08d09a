// Host-side construction of the per-order operator matrices and of the
// subcell geometry, in FP64.  The formulas are the textbook ones the
// reference builds with NumPy (src/polynomials.py:5-122,
// src/sd_simulator.py:44-52, src/sdader_simulator.py:40-92); the Gauss-Legendre
// rule is computed by Newton iteration on P_n and symmetrised, the inverses by
// Gauss-Jordan elimination with partial pivoting.
#pragma once
#include <cmath>
#include <stdexcept>
#include <vector>

namespace spd {
namespace host {

using Mat = std::vector<std::vector<double>>;

// n-point Gauss-Legendre rule on [0, 1], nodes ascending.
inline void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  if (n == 0) return;
  std::vector<double> xs(n), ws(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      double pn = (n == 1) ? z : p1;
      double pm = (n == 1) ? 1.0 : p0;
      dp = n * (z * pn - pm) / (z * z - 1.0);
      double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    // recompute the derivative at the converged node
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = pk;
    }
    double pn = (n == 1) ? z : p1;
    double pm = (n == 1) ? 1.0 : p0;
    dp = n * (z * pn - pm) / (z * z - 1.0);
    xs[n - 1 - i] = z;  // cos-guesses are descending
    ws[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  // symmetrise on [-1, 1], then map to [0, 1]
  for (int i = 0; i < n; ++i) {
    double xi = (xs[i] - xs[n - 1 - i]) / 2.0;
    double wi = (ws[i] + ws[n - 1 - i]) / 2.0;
    x[i] = (xi + 1.0) * 1.0 / 2.0;
    w[i] = wi * 1.0 / 2.0;
  }
}

// Gauss-Chebyshev solution points on [0, 1].
inline std::vector<double> solution_points(int p) {
  std::vector<double> s(p + 1);
  for (int i = 0; i <= p; ++i)
    s[i] = (1.0 - std::cos(double(2 * i + 1) / double(2 * (p + 1)) * M_PI)) * 1.0 / 2.0;
  return s;
}

// {0, GL(p), 1}.
inline std::vector<double> flux_points(int p) {
  std::vector<double> x, w;
  gauss_legendre(p, x, w);
  std::vector<double> f(p + 2);
  f[0] = 0.0;
  for (int i = 0; i < p; ++i) f[i + 1] = x[i];
  f[p + 1] = 1.0;
  return f;
}

// L[a][b] = l_b(x_to[a]), Lagrange basis on x_from.
inline Mat lagrange(const std::vector<double>& to, const std::vector<double>& from) {
  const int m = (int)to.size(), n = (int)from.size();
  Mat L(m, std::vector<double>(n));
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < n; ++b) {
      double num = 1.0, den = 1.0;
      for (int k = 0; k < n; ++k) {
        num *= (k == b) ? 1.0 : (to[a] - from[k]);
        den *= (k == b) ? 1.0 : (from[b] - from[k]);
      }
      L[a][b] = num / den;
    }
  return L;
}

// D[a][b] = l_b'(x_to[a]).
inline Mat lagrange_prime(const std::vector<double>& to, const std::vector<double>& from) {
  const int m = (int)to.size(), n = (int)from.size();
  Mat D(m, std::vector<double>(n, 0.0));
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < n; ++b) {
      double den = 1.0;
      for (int k = 0; k < n; ++k) den *= (k == b) ? 1.0 : (from[b] - from[k]);
      double acc = 0.0;
      for (int l = 0; l < n; ++l) {
        if (l == b) continue;
        double num = 1.0;
        for (int k = 0; k < n; ++k) num *= (k == b || k == l) ? 1.0 : (to[a] - from[k]);
        acc += num / den;
      }
      D[a][b] = acc;
    }
  return D;
}

// M[k][b] = mean of l_b over [x_fp[k], x_fp[k+1]] with a p-point GL rule.
inline Mat subcell_mean(const std::vector<double>& sp, const std::vector<double>& fp) {
  const int n = (int)sp.size() - 1;
  const int ncv = (int)fp.size() - 1;
  Mat M(ncv, std::vector<double>(sp.size(), 0.0));
  if (n == 0) {
    M[0][0] = 1.0;
    return M;
  }
  std::vector<double> xq, wq;
  gauss_legendre(n, xq, wq);
  for (int k = 0; k < ncv; ++k) {
    std::vector<double> pts(n);
    for (int q = 0; q < n; ++q) pts[q] = fp[k] + (fp[k + 1] - fp[k]) * xq[q];
    Mat L = lagrange(pts, sp);
    for (size_t b = 0; b < sp.size(); ++b) {
      double acc = 0.0;
      for (int q = 0; q < n; ++q) acc += L[q][b] * wq[q];
      M[k][b] = acc;
    }
  }
  return M;
}

inline Mat inverse(Mat A) {
  const int n = (int)A.size();
  Mat I(n, std::vector<double>(n, 0.0));
  for (int i = 0; i < n; ++i) I[i][i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    if (A[piv][c] == 0.0) throw std::runtime_error("singular operator matrix");
    std::swap(A[piv], A[c]);
    std::swap(I[piv], I[c]);
    const double d = A[c][c];
    for (int k = 0; k < n; ++k) {
      A[c][k] /= d;
      I[c][k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = A[r][c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        A[r][k] -= f * A[c][k];
        I[r][k] -= f * I[c][k];
      }
    }
  }
  return I;
}

struct OpsHost {
  int p, m;
  std::vector<double> sp, fp, x_tp, w_tp;
  Mat sp_to_fp, dfp_to_sp, sp_to_cv, cv_to_sp, invader;
  Mat sp_to_reg;  // solution points -> p+1 equispaced subcell centres (regular_mesh)

  OpsHost(int p_, int m_) : p(p_), m(m_) {
    sp = solution_points(p);
    fp = flux_points(p);
    sp_to_fp = lagrange(fp, sp);
    dfp_to_sp = lagrange_prime(sp, fp);
    sp_to_cv = subcell_mean(sp, fp);
    {
      // centres of p+1 equal subcells: x = arange(p+2)/(p+1), 0.5 (x[1:] + x[:-1])
      std::vector<double> xr(p + 1);
      for (int i = 0; i <= p; ++i) xr[i] = 0.5 * (double(i + 1) / (p + 1) + double(i) / (p + 1));
      sp_to_reg = lagrange(xr, sp);
    }
    cv_to_sp = inverse(sp_to_cv);
    gauss_legendre(m + 1, x_tp, w_tp);
    const int T = m + 1;
    Mat end = lagrange(std::vector<double>{1.0}, x_tp);
    Mat d = lagrange_prime(x_tp, x_tp);  // d[a][b] = l_b'(x_a)
    Mat K(T, std::vector<double>(T));
    for (int i = 0; i < T; ++i)
      for (int j = 0; j < T; ++j) K[i][j] = end[0][j] * end[0][i] - d[j][i] * w_tp[j];
    Mat inv = inverse(K);
    invader = inv;
    for (int i = 0; i < T; ++i)
      for (int j = 0; j < T; ++j) invader[i][j] = w_tp[j] * inv[i][j];
  }
};

// Ghosted subcell geometry of one dimension (src/sdader_simulator.py:61-92):
// face positions (Ns+5), centres (Ns+4), widths (Ns+4), centre spacings (Ns+3).
struct GeomHost {
  std::vector<double> faces, centers, hfp, hcv;
  // reciprocals used by the kernels: 1/hfp, 1/hcv and 1/(c[k+1] - c[k-1])
  std::vector<double> rhfp, rhcv, rdc;
  GeomHost(int N, double h, const std::vector<double>& fp_local, int ngh) {
    const int n = (int)fp_local.size() - 1;
    const int Ns = N * n;
    faces.assign(Ns + 2 * ngh + 1, 0.0);
    for (int k = 0; k < Ns; ++k) faces[ngh + k] = h * (double(k / n) + fp_local[k % n]);
    faces[ngh + Ns] = h * double(N);
    const int L = (int)faces.size();
    std::vector<double> inner(faces.begin() + ngh + 1, faces.begin() + 2 * ngh + 1);
    for (int k = 0; k < ngh; ++k) faces[k] = -inner[ngh - 1 - k];
    const double last = faces[L - (ngh + 1)];
    for (int k = 0; k < ngh; ++k) faces[L - ngh + k] = last + inner[k];
    centers.resize(L - 1);
    hfp.resize(L - 1);
    for (int k = 0; k < L - 1; ++k) {
      centers[k] = 0.5 * (faces[k + 1] + faces[k]);
      hfp[k] = faces[k + 1] - faces[k];
    }
    hcv.resize(L - 2);
    for (int k = 0; k < L - 2; ++k) hcv[k] = centers[k + 1] - centers[k];
    rhfp.resize(hfp.size());
    for (size_t k = 0; k < hfp.size(); ++k) rhfp[k] = 1.0 / hfp[k];
    rhcv.resize(hcv.size());
    for (size_t k = 0; k < hcv.size(); ++k) rhcv[k] = 1.0 / hcv[k];
    rdc.assign(centers.size(), 0.0);
    for (size_t k = 1; k + 1 < centers.size(); ++k) rdc[k] = 1.0 / (centers[k + 1] - centers[k - 1]);
  }
};

}  // namespace host
}  // namespace spd
