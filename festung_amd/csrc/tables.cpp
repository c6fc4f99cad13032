// Reference-element tables (host side).  Quadrature rules follow
// core/quadRule1D.m:48-95, core/quadRule2D.m:64-109 and core/triquad.m:30-76;
// the basis follows core/phi.m:40-61 and core/gradPhi.m:42-87; the tables
// follow core/computeBasesOnQuad.m:72-134 and the reference blocks
// core/integrateRefElemPhiPhi.m:49-59, core/integrateRefElemDphiPhiPhi.m:55-82,
// core/integrateRefEdgePhiIntPhiIntPerQuad.m:52-66 and
// core/integrateRefEdgePhiIntPhiExtPerQuad.m:54-70.
#include "tables.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

#include "common.hpp"

namespace festung {

Rule1D quad_rule_1d(int qOrd) {
  std::vector<double> Q, W;
  switch (qOrd) {
    case 0: case 1: Q = {0.0}; W = {2.0}; break;
    case 2: case 3: {
      double a = std::sqrt(1.0 / 3);
      Q = {-a, a}; W = {1.0, 1.0}; break;
    }
    case 4: case 5: {
      double a = std::sqrt(3.0 / 5);
      Q = {-a, 0.0 * a, a}; W = {1.0 / 9 * 5, 1.0 / 9 * 8, 1.0 / 9 * 5}; break;
    }
    case 6: case 7: {
      const double sg[4] = {-1, -1, 1, 1}, pm[4] = {1, -1, -1, 1};
      for (int i = 0; i < 4; ++i) Q.push_back(sg[i] * std::sqrt(3.0 / 7 + pm[i] * 2 / 7.0 * std::sqrt(6.0 / 5)));
      const double wm[4] = {-1, 1, 1, -1};
      for (int i = 0; i < 4; ++i) W.push_back(1.0 / 36 * (18 + std::sqrt(30.0) * wm[i]));
      break;
    }
    case 8: case 9: {
      const double sg[5] = {-1, -1, 0, 1, 1}, pm[5] = {2, -2, 0, -2, 2};
      for (int i = 0; i < 5; ++i) Q.push_back(sg[i] * std::sqrt(5 + pm[i] * std::sqrt(10.0 / 7)) / 3);
      const double wm[5] = {-1, 1, 0, 1, -1}, add[5] = {0, 0, 190, 0, 0};
      for (int i = 0; i < 5; ++i) W.push_back(1.0 / 900 * (322 + 13 * std::sqrt(70.0) * wm[i] + add[i]));
      break;
    }
    case 10: case 11:
      Q = {0.6612093864662645, -0.6612093864662645, -0.2386191860831969,
           0.2386191860831969, -0.9324695142031521, 0.9324695142031521};
      W = {0.3607615730481386, 0.3607615730481386, 0.4679139345726910,
           0.4679139345726910, 0.1713244923791704, 0.171324492379170};
      break;
    case 12: case 13:
      Q = {0.0000000000000000, 0.4058451513773972, -0.4058451513773972,
           -0.7415311855993945, 0.7415311855993945, -0.9491079123427585,
           0.9491079123427585};
      W = {0.4179591836734694, 0.3818300505051189, 0.3818300505051189,
           0.2797053914892766, 0.2797053914892766, 0.1294849661688697,
           0.1294849661688697};
      break;
    case 14: case 15:
      Q = {-0.1834346424956498, 0.1834346424956498, -0.5255324099163290,
           0.5255324099163290, -0.7966664774136267, 0.7966664774136267,
           -0.9602898564975363, 0.9602898564975363};
      W = {0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
           0.3137066458778873, 0.2223810344533745, 0.2223810344533745,
           0.1012285362903763, 0.1012285362903763};
      break;
    case 16: case 17:
      Q = {0.0000000000000000, -0.8360311073266358, 0.8360311073266358,
           -0.9681602395076261, 0.9681602395076261, -0.3242534234038089,
           0.3242534234038089, -0.6133714327005904, 0.6133714327005904};
      W = {0.3302393550012598, 0.1806481606948574, 0.1806481606948574,
           0.0812743883615744, 0.0812743883615744, 0.3123470770400029,
           0.3123470770400029, 0.2606106964029354, 0.2606106964029354};
      break;
    default: fail(FESTUNG_ERR_ARG, "quadRule1D: order %d not implemented", qOrd);
  }
  Rule1D r;
  for (size_t i = 0; i < Q.size(); ++i) {
    r.q.push_back((Q[i] + 1) / 2);
    r.w.push_back(W[i] / 2);
  }
  return r;
}

// Eigen-decomposition of a symmetric tridiagonal matrix by the implicit QL
// method (Bowdler, Martin, Reinsch & Wilkinson, "The QR and QL algorithms for
// symmetric matrices", Numer. Math. 11, 1968; EISPACK tql2).  d: diagonal
// (overwritten by eigenvalues), e: sub-diagonal in e[1..n-1]; z: n x n
// column-major, on exit the eigenvectors.
static void tql2(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& z) {
  z.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) z[i + i * n] = 1.0;
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  double f = 0.0, tst1 = 0.0;
  for (int l = 0; l < n; ++l) {
    tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
    int m = l;
    while (m < n - 1) {
      if (std::fabs(e[m]) <= DBL_EPSILON * tst1) break;
      ++m;
    }
    if (m > l) {
      for (int iter = 0; iter < 60; ++iter) {
        double g = d[l];
        double p = (d[l + 1] - g) / (2.0 * e[l]);
        double r = std::hypot(p, 1.0);
        if (p < 0) r = -r;
        d[l] = e[l] / (p + r);
        d[l + 1] = e[l] * (p + r);
        double dl1 = d[l + 1];
        double h = g - d[l];
        for (int i = l + 2; i < n; ++i) d[i] -= h;
        f += h;
        p = d[m];
        double c = 1.0, c2 = c, c3 = c, el1 = e[l + 1], s = 0.0, s2 = 0.0;
        for (int i = m - 1; i >= l; --i) {
          c3 = c2; c2 = c; s2 = s;
          g = c * e[i];
          h = c * p;
          r = std::hypot(p, e[i]);
          e[i + 1] = s * r;
          s = e[i] / r;
          c = p / r;
          p = c * d[i] - s * g;
          d[i + 1] = h + s * (c * g + s * d[i]);
          for (int k = 0; k < n; ++k) {
            h = z[k + (i + 1) * n];
            z[k + (i + 1) * n] = s * z[k + i * n] + c * h;
            z[k + i * n] = c * z[k + i * n] - s * h;
          }
        }
        p = -s * s2 * c3 * el1 * e[l] / dl1;
        e[l] = s * p;
        d[l] = c * p;
        if (std::fabs(e[l]) <= DBL_EPSILON * tst1) break;
      }
    }
    d[l] = d[l] + f;
    e[l] = 0.0;
  }
}

void triquad(int N, std::vector<double>& X, std::vector<double>& Y,
             std::vector<double>& Wx, std::vector<double>& Wy) {
  // Gauss-Jacobi (alpha = 0, beta = 1) by Golub-Welsch, core/triquad.m:56-60
  std::vector<double> A(N + 1), nnk(N);
  A[0] = 1.0 / 3;
  for (int n = 1; n <= N; ++n) {
    nnk[n - 1] = 2.0 * n + 1;
    A[n] = 1.0 / (nnk[n - 1] * (nnk[n - 1] + 2));
  }
  std::vector<double> ab2(N + 1);
  ab2[0] = 2.0;
  ab2[1] = 2.0 / 9;
  for (int n = 2; n <= N; ++n) {
    double k = nnk[n - 1], nk = n + 1.0, k2 = k * k;
    double Bn = 4 * (n * nk) * (n * nk) / (k2 * k2 - k2);
    ab2[n] = Bn;
  }
  std::vector<double> d(A.begin(), A.begin() + N), e(N, 0.0), z;
  for (int i = 1; i < N; ++i) e[i] = std::sqrt(ab2[i]);
  tql2(N, d, e, z);
  std::vector<int> I(N);
  std::iota(I.begin(), I.end(), 0);
  std::stable_sort(I.begin(), I.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> x(N), wx(N);
  for (int i = 0; i < N; ++i) {
    x[i] = (d[I[i]] + 1) / 2;
    double v = z[0 + I[i] * N];
    wx[i] = ab2[0] * v * v / 4;
  }
  // Gauss-Legendre by Newton iteration, core/triquad.m:62-71
  int Nm = N - 1, N1 = Nm + 1, N2 = Nm + 2;
  std::vector<double> y(N1), y0(N1, 2.0), Lp(N1), L(size_t(N1) * N2);
  for (int i = 0; i < N1; ++i) y[i] = std::cos((2.0 * (Nm - i) + 1) * M_PI / (2.0 * Nm + 2));
  auto maxdiff = [&]() {
    double m = 0;
    for (int i = 0; i < N1; ++i) m = std::max(m, std::fabs(y[i] - y0[i]));
    return m;
  };
  while (maxdiff() > DBL_EPSILON) {
    for (int i = 0; i < N1; ++i) {
      L[i + 0 * N1] = 1;
      L[i + 1 * N1] = y[i];
    }
    for (int k = 2; k <= N1; ++k)
      for (int i = 0; i < N1; ++i)
        L[i + k * N1] = ((2 * k - 1) * y[i] * L[i + (k - 1) * N1] - (k - 1) * L[i + (k - 2) * N1]) / k;
    for (int i = 0; i < N1; ++i) {
      Lp[i] = N2 * (L[i + (N1 - 1) * N1] - y[i] * L[i + (N2 - 1) * N1]) / (1 - y[i] * y[i]);
      y0[i] = y[i];
      y[i] = y0[i] - L[i + (N2 - 1) * N1] / Lp[i];
    }
  }
  // collapse onto the triangle v = [0 0; 0 1; 1 0], core/triquad.m:73-76
  Wx.resize(N);
  Wy.resize(N1);
  for (int i = 0; i < N; ++i) Wx[i] = 1.0 * wx[i];  // |det(cd(2:3,:))| = 1
  for (int j = 0; j < N1; ++j)
    Wy[j] = 1 / ((1 - y[j] * y[j]) * Lp[j] * Lp[j]) * ((double(N2) / N1) * (double(N2) / N1));
  X.assign(size_t(N) * N1, 0);
  Y.assign(size_t(N) * N1, 0);
  for (int j = 0; j < N1; ++j) {
    double t = (1 + y[j]) / 2;
    for (int i = 0; i < N; ++i) {
      double xx = x[i], yy = t * xx;
      X[i + j * N] = 0.0 + 1.0 * xx + (-1.0) * yy;
      Y[i + j * N] = 0.0 + 0.0 * xx + 1.0 * yy;
    }
  }
}

Rule2D quad_rule_2d(int qOrd) {
  Rule2D r;
  switch (qOrd) {
    case 0: case 1: r.q1 = {1.0 / 3}; r.q2 = {1.0 / 3}; r.w = {1.0 / 2}; break;
    case 2:
      r.q1 = {1.0 / 6, 2.0 / 3, 1.0 / 6}; r.q2 = {1.0 / 6, 1.0 / 6, 2.0 / 3};
      r.w = {1.0 / 6, 1.0 / 6, 1.0 / 6};
      break;
    case 3: {
      double a1 = (3 + std::sqrt(3.0)) / 6, a2 = (3 - std::sqrt(3.0)) / 6;
      double b1 = (6 + std::sqrt(6.0)) / 10, b2 = (6 - std::sqrt(6.0)) / 10;
      double w1 = 1.0 / 8 + std::sqrt(6.0) / 72, w2 = 1.0 / 8 - std::sqrt(6.0) / 72;
      r.q1 = {a1 * b1, a2 * b1, a1 * b2, a2 * b2};
      r.q2 = {a2 * b1, a1 * b1, a2 * b2, a1 * b2};
      r.w = {w1, w1, w2, w2};
      break;
    }
    case 4:
      r.q1 = {0.445948490915965, 0.108103018168070, 0.445948490915965,
              0.091576213509771, 0.816847572980458, 0.091576213509771};
      r.q2 = {0.108103018168070, 0.445948490915965, 0.445948490915965,
              0.816847572980458, 0.091576213509771, 0.091576213509771};
      r.w = {0.111690794839005, 0.111690794839005, 0.111690794839005,
             0.054975871827661, 0.054975871827661, 0.054975871827661};
      break;
    case 5: {
      double s15 = std::sqrt(15.0);
      double a1 = (6 - s15) / 21, a2 = (6 + s15) / 21;
      double w1 = (155 - s15) / 2400, w2 = (155 + s15) / 2400;
      r.q1 = {1.0 / 3, a1, 1 - 2 * a1, a1, a2, 1 - 2 * a2, a2};
      r.q2 = {1.0 / 3, 1 - 2 * a1, a1, a1, 1 - 2 * a2, a2, a2};
      r.w = {9.0 / 80, w1, w1, w1, w2, w2, w2};
      break;
    }
    case 6:
      r.q1 = {0.063089014491502, 0.873821971016996, 0.063089014491502,
              0.249286745170910, 0.501426509658179, 0.249286745170910,
              0.310352451033785, 0.053145049844816, 0.636502499121399,
              0.053145049844816, 0.636502499121399, 0.310352451033785};
      r.q2 = {0.063089014491502, 0.063089014491502, 0.873821971016996,
              0.249286745170910, 0.249286745170910, 0.501426509658179,
              0.053145049844816, 0.310352451033785, 0.053145049844816,
              0.636502499121399, 0.310352451033785, 0.636502499121399};
      r.w = {0.025422453185103, 0.025422453185103, 0.025422453185103,
             0.058393137863189, 0.058393137863189, 0.058393137863189,
             0.041425537809187, 0.041425537809187, 0.041425537809187,
             0.041425537809187, 0.041425537809187, 0.041425537809187};
      break;
    default: {
      if (qOrd < 0 || qOrd > 17) fail(FESTUNG_ERR_ARG, "quadRule2D: order %d not implemented", qOrd);
      std::vector<double> X, Y, Wx, Wy;
      triquad(qOrd, X, Y, Wx, Wy);
      r.q1 = X;
      r.q2 = Y;
      r.w.resize(X.size());
      for (int j = 0; j < qOrd; ++j)
        for (int i = 0; i < qOrd; ++i) r.w[i + j * qOrd] = Wx[i] * Wy[j];
    }
  }
  return r;
}

// One definition of the basis in Horner form, instantiated for double
// (evaluation) and for PolyT (monomial coefficients: in double, and in
// __float128 for the Taylor change of basis that the double-double limiter
// path reads as hi + lo pairs).
template <class S>
struct PolyT {
  using scalar = S;
  S c[5][5] = {};
  PolyT() = default;
  explicit PolyT(S v) { c[0][0] = v; }
};
using Poly = PolyT<double>;
template <class S>
using PS = typename PolyT<S>::scalar;  // a scalar operand: S is deduced from the polynomial
template <class S>
static PolyT<S> operator+(const PolyT<S>& a, const PolyT<S>& b) {
  PolyT<S> r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) r.c[i][j] = a.c[i][j] + b.c[i][j];
  return r;
}
template <class S>
static PolyT<S> operator-(const PolyT<S>& a) {
  PolyT<S> r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) r.c[i][j] = -a.c[i][j];
  return r;
}
template <class S>
static PolyT<S> operator-(const PolyT<S>& a, const PolyT<S>& b) { return a + (-b); }
template <class S>
static PolyT<S> operator*(const PolyT<S>& a, const PolyT<S>& b) {
  PolyT<S> r;
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j)
      if (a.c[i][j] != 0)
        for (int k = 0; i + k < 5; ++k)
          for (int l = 0; j + l < 5; ++l) r.c[i + k][j + l] += a.c[i][j] * b.c[k][l];
  return r;
}
template <class S>
static PolyT<S> operator+(PS<S> a, const PolyT<S>& b) { return PolyT<S>(a) + b; }
template <class S>
static PolyT<S> operator+(const PolyT<S>& a, PS<S> b) { return a + PolyT<S>(b); }
template <class S>
static PolyT<S> operator-(PS<S> a, const PolyT<S>& b) { return PolyT<S>(a) - b; }
template <class S>
static PolyT<S> operator-(const PolyT<S>& a, PS<S> b) { return a - PolyT<S>(b); }
template <class S>
static PolyT<S> operator*(PS<S> a, const PolyT<S>& b) { return PolyT<S>(a) * b; }

static double sqrt_s(double x) { return std::sqrt(x); }
static __float128 sqrt_s(__float128 x) {
  // Newton's iteration r <- (r + x / r) / 2 from the double root: 53 -> 106 -> 212 bits
  __float128 r = std::sqrt(double(x));
  r = (r + x / r) / 2;
  return (r + x / r) / 2;
}

template <class S, class T>
static T phi_t(int i, const T& X1, const T& X2) {
  const S S2 = sqrt_s(S(2)), S3 = sqrt_s(S(3)), S5 = sqrt_s(S(5)), S6 = sqrt_s(S(6));
  const S S10 = sqrt_s(S(10)), S14 = sqrt_s(S(14)), S30 = sqrt_s(S(30)), S70 = sqrt_s(S(70));
  switch (i) {
    case 1: return S2 * (0.0 * X1 + 1.0);
    case 2: return 2 - 6 * X1;
    case 3: return 2 * S3 * (1 - X1 - 2 * X2);
    case 4: return S6 * ((10 * X1 - 8) * X1 + 1);
    case 5: return S3 * ((5 * X1 - 4) * X1 + (-15 * X2 + 12) * X2 - 1);
    case 6: return 3 * S5 * ((3 * X1 + 8 * X2 - 4) * X1 + (3 * X2 - 4) * X2 + 1);
    case 7: return 2 * S2 * (-1 + (15 + (-45 + 35 * X1) * X1) * X1);
    case 8: return 2 * S6 * (-1 + (13 + (-33 + 21 * X1) * X1) * X1 + (2 + (-24 + 42 * X1) * X1) * X2);
    case 9: return 2 * S10 * (-1 + (9 + (-15 + 7 * X1) * X1) * X1 + (6 + (-48 + 42 * X1) * X1 + (-6 + 42 * X1) * X2) * X2);
    case 10: return 2 * S14 * (-1 + (3 + (-3 + X1) * X1) * X1 + (12 + (-24 + 12 * X1) * X1 + (-30 + 30 * X1 + 20 * X2) * X2) * X2);
    case 11: return S10 * (1 + (-24 + (126 + (-224 + 126 * X1) * X1) * X1) * X1);
    case 12: return S30 * (1 + (-22 + (105 + (-168 + 84 * X1) * X1) * X1) * X1 + (-2 + (42 + (-168 + 168 * X1) * X1) * X1) * X2);
    case 13: return 5 * S2 * (1 + (-18 + (69 + (-88 + 36 * X1) * X1) * X1) * X1 + (-6 + (102 + (-312 + 216 * X1) * X1) * X1
                              + (6 + (-96 + 216 * X1) * X1) * X2) * X2);
    case 14: return S70 * (1 + (-12 + (30 + (-28 + 9 * X1) * X1) * X1) * X1 + (-12 + (132 + (-228 + 108 * X1) * X1) * X1
                           + (30 + (-300 + 270 * X1) * X1 + (-20 + 180 * X1) * X2) * X2) * X2);
    case 15: return 3 * S10 * (1 + (-4 + (6 + (-4 + X1) * X1) * X1) * X1 + (-20 + (60 + (-60 + 20 * X1) * X1) * X1
                               + (90 + (-180 + 90 * X1) * X1 + (-140 + 140 * X1 + 70 * X2) * X2) * X2) * X2);
  }
  fail(FESTUNG_ERR_ARG, "phi: index %d out of range", i);
}

double phi(int i, double x1, double x2) {
  if (i == 1) return std::sqrt(2.0);  // sqrt(2)*ones(size(X1))
  return phi_t<double>(i, x1, x2);
}

double grad_phi(int i, int m, double X1, double X2) {
  const double S2 = std::sqrt(2.0), S3 = std::sqrt(3.0), S5 = std::sqrt(5.0), S6 = std::sqrt(6.0);
  const double S10 = std::sqrt(10.0), S14 = std::sqrt(14.0), S30 = std::sqrt(30.0), S70 = std::sqrt(70.0);
  if (m == 1) {
    switch (i) {
      case 1: return 0.0;
      case 2: return -6.0;
      case 3: return -2 * S3;
      case 4: return S6 * (20 * X1 - 8);
      case 5: return S3 * (10 * X1 - 4);
      case 6: return 6 * S5 * (3 * X1 + 4 * X2 - 2);
      case 7: return 2 * S2 * (15 + (-90 + 105 * X1) * X1);
      case 8: return 2 * S6 * (13 + (-66 + 63 * X1) * X1 + (-24 + 84 * X1) * X2);
      case 9: return 2 * S10 * (9 + (-30 + 21 * X1) * X1 + (-48 + 84 * X1 + 42 * X2) * X2);
      case 10: return 2 * S14 * (3 + (-6 + 3 * X1) * X1 + (-24 + 24 * X1 + 30 * X2) * X2);
      case 11: return S10 * (-24 + (252 + (-672 + 504 * X1) * X1) * X1);
      case 12: return S30 * (-22 + (210 + (-504 + 336 * X1) * X1) * X1 + (42 + (-336 + 504 * X1) * X1) * X2);
      case 13: return 5 * S2 * (-18 + (138 + (-264 + 144 * X1) * X1) * X1
                                + (102 + (-624 + 648 * X1) * X1 + (-96 + 432 * X1) * X2) * X2);
      case 14: return S70 * (-12 + (60 + (-84 + 36 * X1) * X1) * X1
                             + (132 + (-456 + 324 * X1) * X1 + (-300 + 540 * X1 + 180 * X2) * X2) * X2);
      case 15: return 3 * S10 * (-4 + (12 + (-12 + 4 * X1) * X1) * X1
                                 + (60 + (-120 + 60 * X1) * X1 + (-180 + 180 * X1 + 140 * X2) * X2) * X2);
    }
  } else {
    switch (i) {
      case 1: case 2: case 4: case 7: case 11: return 0.0;
      case 3: return -4 * S3;
      case 5: return 2 * S3 * (-15 * X2 + 6);
      case 6: return 6 * S5 * (4 * X1 + 3 * X2 - 2);
      case 8: return 2 * S6 * (2 + (-24 + 42 * X1) * X1);
      case 9: return 2 * S10 * (6 + (-48 + 42 * X1) * X1 + (-12 + 84 * X1) * X2);
      case 10: return 2 * S14 * (12 + (-24 + 12 * X1) * X1 + (-60 + 60 * X1 + 60 * X2) * X2);
      case 12: return S30 * (-2 + (42 + (-168 + 168 * X1) * X1) * X1);
      case 13: return 5 * S2 * (-6 + (102 + (-312 + 216 * X1) * X1) * X1
                                + (12 + (-192 + 432 * X1) * X1) * X2);
      case 14: return S70 * (-12 + (132 + (-228 + 108 * X1) * X1) * X1
                             + (60 + (-600 + 540 * X1) * X1 + (-60 + 540 * X1) * X2) * X2);
      case 15: return 3 * S10 * (-20 + (60 + (-60 + 20 * X1) * X1) * X1
                                 + (180 + (-360 + 180 * X1) * X1 + (-420 + 420 * X1 + 280 * X2) * X2) * X2);
    }
  }
  fail(FESTUNG_ERR_ARG, "gradPhi: index (%d, %d) out of range", i, m);
}

void gamma_map(int n, double s, double& x1, double& x2) {
  switch (n) {
    case 1: x1 = 1 - s; x2 = s; return;
    case 2: x1 = 0; x2 = 1 - s; return;
    case 3: x1 = s; x2 = 0; return;
  }
  fail(FESTUNG_ERR_ARG, "gammaMap: edge %d", n);
}

void theta_map(int nn, int np, double X1, double X2, double& y1, double& y2) {
  switch (nn * 4 + np) {
    case 1 * 4 + 1: y1 = 1 - X1; y2 = 1 - X2; return;
    case 1 * 4 + 2: y1 = 0; y2 = X2; return;
    case 1 * 4 + 3: y1 = X1; y2 = 0; return;
    case 2 * 4 + 1: y1 = 1 - X2; y2 = X2; return;
    case 2 * 4 + 2: y1 = 0; y2 = 1 - X2; return;
    case 2 * 4 + 3: y1 = X2; y2 = 0; return;
    case 3 * 4 + 1: y1 = X1; y2 = 1 - X1; return;
    case 3 * 4 + 2: y1 = 0; y2 = X1; return;
    case 3 * 4 + 3: y1 = 1 - X1; y2 = 0; return;
  }
  fail(FESTUNG_ERR_ARG, "theta: (%d, %d)", nn, np);
}

template <class S>
static S abs_s(S x) { return x < 0 ? -x : x; }

template <class S>
static std::vector<S> invert_t(const std::vector<S>& A0, int n) {
  std::vector<S> A(A0), X(size_t(n) * n, S(0));
  for (int i = 0; i < n; ++i) X[i + i * n] = 1;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (abs_s(A[r + c * n]) > abs_s(A[piv + c * n])) piv = r;
    if (A[piv + c * n] == 0) fail(FESTUNG_ERR_INTERNAL, "singular matrix");
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(A[c + k * n], A[piv + k * n]);
        std::swap(X[c + k * n], X[piv + k * n]);
      }
    S d = A[c + c * n];
    for (int k = 0; k < n; ++k) {
      A[c + k * n] /= d;
      X[c + k * n] /= d;
    }
    for (int r = 0; r < n; ++r)
      if (r != c) {
        S f = A[r + c * n];
        if (f != 0)
          for (int k = 0; k < n; ++k) {
            A[r + k * n] -= f * A[c + k * n];
            X[r + k * n] -= f * X[c + k * n];
          }
      }
  }
  return X;
}

std::vector<double> invert(const std::vector<double>& A0, int n) { return invert_t(A0, n); }

static double fact(int n) {
  double r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

RefTables build_ref_tables(int p) {
  if (p < 0 || p > 4) fail(FESTUNG_ERR_UNSUPPORTED, "polynomial degree p = %d (must be 0..4)", p);
  RefTables t;
  t.p = p;
  const int N = t.N = nbasis(p);
  t.ordE = std::max(2 * p, 1);
  t.ordF = 2 * p + 1;
  t.ruleE = quad_rule_2d(t.ordE);
  t.ruleF = quad_rule_2d(t.ordF);
  t.ruleEdge = quad_rule_1d(2 * p + 1);
  t.RE = int(t.ruleE.w.size());
  t.RF = int(t.ruleF.w.size());
  t.Re = int(t.ruleEdge.w.size());
  const int RE = t.RE, RF = t.RF, Re = t.Re;
  t.phiE.resize(size_t(RE) * N);
  t.gradE.resize(size_t(RE) * N * 2);
  for (int i = 0; i < N; ++i)
    for (int q = 0; q < RE; ++q) {
      t.phiE[q + i * RE] = phi(i + 1, t.ruleE.q1[q], t.ruleE.q2[q]);
      for (int m = 0; m < 2; ++m)
        t.gradE[q + i * RE + m * RE * N] = grad_phi(i + 1, m + 1, t.ruleE.q1[q], t.ruleE.q2[q]);
    }
  if (p == 4) {
    static const double kRule11[28][3] = {
#include "tri_rule11.inc"
    };
    for (const auto& r : kRule11) {
      t.ruleG.q1.push_back(r[0]);
      t.ruleG.q2.push_back(r[1]);
      t.ruleG.w.push_back(r[2]);
    }
  } else {
    t.ruleG = t.ruleE;
  }
  const int RG = t.RG = int(t.ruleG.w.size());
  t.phiG.resize(size_t(RG) * N);
  t.gradG.resize(size_t(RG) * N * 2);
  for (int i = 0; i < N; ++i)
    for (int q = 0; q < RG; ++q) {
      t.phiG[q + i * RG] = phi(i + 1, t.ruleG.q1[q], t.ruleG.q2[q]);
      for (int m = 0; m < 2; ++m)
        t.gradG[q + i * RG + m * RG * N] = grad_phi(i + 1, m + 1, t.ruleG.q1[q], t.ruleG.q2[q]);
    }
  t.phiF.resize(size_t(RF) * N);
  for (int i = 0; i < N; ++i)
    for (int q = 0; q < RF; ++q) t.phiF[q + i * RF] = phi(i + 1, t.ruleF.q1[q], t.ruleF.q2[q]);
  t.phi1D.resize(size_t(Re) * N * 3);
  t.thetaPhi1D.resize(size_t(Re) * N * 9);
  for (int nn = 0; nn < 3; ++nn)
    for (int r = 0; r < Re; ++r) {
      double a1, a2;
      gamma_map(nn + 1, t.ruleEdge.q[r], a1, a2);
      for (int i = 0; i < N; ++i) t.phi1D[r + i * Re + nn * Re * N] = phi(i + 1, a1, a2);
      for (int np = 0; np < 3; ++np) {
        double b1, b2;
        theta_map(nn + 1, np + 1, a1, a2, b1, b2);
        for (int i = 0; i < N; ++i)
          t.thetaPhi1D[r + i * Re + nn * Re * N + np * Re * N * 3] = phi(i + 1, b1, b2);
      }
    }
  // M-hat, core/integrateRefElemPhiPhi.m:49-59 (order max(2p, 1) == ordE)
  t.hatM.assign(size_t(N) * N, 0.0);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double s = 0;
      for (int q = 0; q < RE; ++q) s += t.ruleE.w[q] * t.phiE[q + i * RE] * t.phiE[q + j * RE];
      t.hatM[i + j * N] = s;
    }
  t.hatMinv = invert(t.hatM, N);
  // G-hat, core/integrateRefElemDphiPhiPhi.m:55-82
  t.hatG.assign(size_t(N) * N * N * 2, 0.0);
  if (N > 1)
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j)
        for (int l = 0; l < N; ++l)
          for (int m = 0; m < 2; ++m) {
            double s = 0;
            for (int q = 0; q < RE; ++q)
              s += t.ruleE.w[q] * t.gradE[q + i * RE + m * RE * N] * t.phiE[q + j * RE] * t.phiE[q + l * RE];
            t.hatG[i + N * (j + N * (l + N * m))] = s;
          }
  // R-hat blocks, core/integrateRefEdgePhiIntPhiIntPerQuad.m:52-66 and
  // core/integrateRefEdgePhiIntPhiExtPerQuad.m:54-70
  t.hatRdiag.assign(size_t(N) * N * 3 * Re, 0.0);
  t.hatRoff.assign(size_t(N) * N * 9 * Re, 0.0);
  for (int r = 0; r < Re; ++r)
    for (int nn = 0; nn < 3; ++nn)
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          double pi = t.phi1D[r + i * Re + nn * Re * N];
          t.hatRdiag[i + N * (j + N * (nn + 3 * r))] = pi * t.phi1D[r + j * Re + nn * Re * N] * t.ruleEdge.w[r];
          for (int np = 0; np < 3; ++np)
            t.hatRoff[i + N * (j + N * (nn + 3 * (np + 3 * r)))] =
                pi * t.thetaPhi1D[r + j * Re + nn * Re * N + np * Re * N * 3] * t.ruleEdge.w[r];
        }
  // Taylor change of basis: row 0 = mean (sqrt(2) e_1 by orthonormality),
  // row idx(a, b) = d1^a d2^b phi_j at the centroid (1/3, 1/3), Taylor order
  // (1,0),(0,1),(2,0),(1,1),(0,2),...  Computed in __float128 and kept as
  // the nearest double (F, Finv) plus the remainder (Flo, Finvlo).
  using Q = __float128;
  std::vector<Q> Fq(size_t(N) * N, Q(0));
  Fq[0] = sqrt_s(Q(2));
  for (int j = 0; j < N; ++j) {
    PolyT<Q> x1, x2;
    x1.c[1][0] = 1;
    x2.c[0][1] = 1;
    const PolyT<Q> ph = phi_t<Q>(j + 1, x1, x2);
    for (int d = 1; d <= p; ++d)
      for (int b = 0; b <= d; ++b) {
        int a = d - b, row = d * (d + 1) / 2 + b;
        Q s = 0;
        for (int m = a; m < 5; ++m)
          for (int n = b; m + n < 5; ++n) {
            const Q c = ph.c[m][n];
            if (c == 0) continue;
            Q third = 1;
            for (int k = 0; k < m - a + n - b; ++k) third /= 3;
            s += c * Q(fact(m) / fact(m - a)) * Q(fact(n) / fact(n - b)) * third;
          }
        Fq[row + j * N] = s;
      }
  }
  std::vector<Q> Fiq = invert_t(Fq, N);
  // entries that vanish exactly keep a residue of about 1e-35 relative to
  // the largest (the 1/3 powers and the elimination round in __float128)
  for (std::vector<Q>* m : {&Fq, &Fiq}) {
    Q big = 0;
    for (Q v : *m) big = std::max(big, abs_s(v));
    for (Q& v : *m)
      if (abs_s(v) <= big * Q(1e-24)) v = 0;
  }
  t.F.resize(size_t(N) * N);
  t.Flo.resize(size_t(N) * N);
  t.Finv.resize(size_t(N) * N);
  t.Finvlo.resize(size_t(N) * N);
  for (size_t i = 0; i < size_t(N) * N; ++i) {
    t.F[i] = double(Fq[i]);
    t.Flo[i] = double(Fq[i] - Q(t.F[i]));
    t.Finv[i] = double(Fiq[i]);
    t.Finvlo[i] = double(Fiq[i] - Q(t.Finv[i]));
  }
  return t;
}

}  // namespace festung
