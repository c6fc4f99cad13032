// Reference-element tables: quadrature rules, the orthonormal Dubiner basis,
// edge maps and the reference blocks, plus the constant tables of the
// matrix-free Taylor-basis change used by the slope limiter.
#pragma once
#include <vector>

namespace festung {

struct Rule1D { std::vector<double> q, w; };           // nodes / weights on [0, 1]
struct Rule2D { std::vector<double> q1, q2, w; };      // on the reference triangle

// core/quadRule1D.m:48-95
Rule1D quad_rule_1d(int qOrd);
// core/quadRule2D.m:64-109 (with core/triquad.m:30-76 for qOrd >= 7)
Rule2D quad_rule_2d(int qOrd);
// core/triquad.m:30-76 (v = [0 0; 0 1; 1 0]): X, Y column-major N x N, Wx, Wy
void triquad(int n, std::vector<double>& X, std::vector<double>& Y,
             std::vector<double>& Wx, std::vector<double>& Wy);

// core/phi.m:40-61 and core/gradPhi.m:42-87 (i, m are 1-based)
double phi(int i, double x1, double x2);
double grad_phi(int i, int m, double x1, double x2);
// core/gammaMap.m:50-57, core/theta.m:109-133 (nn, np 1-based)
void gamma_map(int n, double s, double& x1, double& x2);
void theta_map(int nn, int np, double x1, double x2, double& y1, double& y2);

// Monomial coefficients of phi_i: coef[a*5 + b] multiplies x1^a x2^b.

// Reference blocks and basis tables for polynomial degree p (all
// column-major, Matlab index order).
struct RefTables {
  int p = 0, N = 0;
  int ordE = 0, ordF = 0;          // element rules 2p and 2p+1 (1 for p = 0)
  Rule2D ruleE, ruleF;             // quadRule2D(ordE), quadRule2D(ordF)
  Rule1D ruleEdge;                 // quadRule1D(2p+1)
  int RE = 0, RF = 0, Re = 0;      // numbers of points
  std::vector<double> phiE, gradE; // phi2D{ordE} (RE x N), gradPhi2D{ordE} (RE x N x 2)
  // Rule of the element term's integrals grad(phi_i) c_h u_h (degree 3p - 1):
  // quadRule2D(ordE) for p <= 3, where that rule is not exact for them and
  // the reference's rounding of the integrals is kept; at p = 4, where
  // quadRule2D(8) integrates them exactly, the 28-point rule of
  // tri_rule11.inc (exact to degree 11), which gives the same integrals.
  Rule2D ruleG;
  int RG = 0;
  std::vector<double> phiG, gradG; // at ruleG's points (RG x N, RG x N x 2)
  std::vector<double> phiF;        // phi2D{ordF} (RF x N)
  std::vector<double> phi1D;       // phi1D{2p+1} (Re x N x 3)
  std::vector<double> thetaPhi1D;  // thetaPhi1D{2p+1} (Re x N x 3 x 3)
  std::vector<double> hatM;        // N x N, core/integrateRefElemPhiPhi.m
  std::vector<double> hatMinv;     // inverse of hatM
  std::vector<double> hatG;        // N x N x N x 2, core/integrateRefElemDphiPhiPhi.m
  std::vector<double> hatRdiag;    // N x N x 3 x Re, core/integrateRefEdgePhiIntPhiIntPerQuad.m
  std::vector<double> hatRoff;     // N x N x 3 x 3 x Re, core/integrateRefEdgePhiIntPhiExtPerQuad.m
  // Taylor change of basis (limiter): F (N x N) maps DG coefficients to
  // [mean; reference-coordinate derivatives at the centroid], Finv its
  // inverse, each the nearest double to the exact value; Flo, Finvlo the
  // remainders (hi + lo is a double-double table).
  std::vector<double> F, Finv, Flo, Finvlo;
};

RefTables build_ref_tables(int p);

// Dense inverse by Gauss-Jordan with partial pivoting (small N).
std::vector<double> invert(const std::vector<double>& A, int n);

}  // namespace festung
