// Reference-structured CPU restatement of one explicit DG advection step
// (test infrastructure: the timed CPU baseline, and a second oracle path
// checked against the numpy oracle).  Every sub-step does what
// advection/preprocessSubStep.m:56-99 and advection/solveSubStep.m:57-99
// do, in their order and data structures:
//   * L2 projections of f, u1, u2 (core/projectFuncCont2DataDisc.m:94-102,
//     point values given) and u . nu at the edge points
//     (core/computeFuncContNuOnQuadEdge.m:46-67, point values given);
//   * G1, G2 assembled as block-diagonal CSR matrices
//     (core/assembleMatElemDphiPhiFuncDisc.m:181-202), R as a CSR matrix of
//     diagonal and neighbour blocks (core/assembleMatEdgePhiPhiValUpwind.m:
//     261-299), the Dirichlet vector K_D (core/assembleVecEdgePhiIntFuncContVal.m:
//     162-185), L = M f;
//   * sysA = -G1 - G2 + R as a new CSR matrix (a sparse sum), one SpMV,
//     and globM \ (.) as a factorisation and solve of every diagonal block
//     (solveSubStep.m:62-66);
//   * with the limiter: the Taylor change of basis as block solves
//     (core/projectDataDisc2DataTaylor.m, core/projectDataTaylor2DataDisc.m),
//     the vertex-patch bounds as row min / max over CSR patch matrices with
//     the global shifts (core/computeMinMaxV0TElementPatch.m:66-112), alpha
//     (core/computeVertexBasedLimiter.m:78-106), the 'linear' limiter
//     (core/applySlopeLimiterTaylorLinear.m:108-122), dc/dt limiting with
//     the mass-lumping correction (solveSubStep.m:69-79);
//   * the SSP RK combination (core/rungeKuttaExplicit.m:64-78, solveSubStep.m:83)
//     or the accumulator form of classic RK4.
// OpenMP over rows / elements; the thread count is an argument.
// Neumann data are not handled (K_N is identically zero on the workloads
// this times; a problem with Neumann marks is rejected).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <vector>

extern "C" {

struct refcpu_input {
  int K, N, R, R2, ordRK, limiter;
  // reference blocks (C order): hatM[N][N], hatG[N][N][N][2], hatRdiag[N][N][3][R],
  // hatRoff[N][N][3][3][R]; element rule of order 2p: W2[R2], phi2[R2][N];
  // edge rule of order 2p+1: W1[R], phi1[R][N][3]
  const double *hatM, *hatG, *hatRdiag, *hatRoff, *W2, *phi2, *W1, *phi1;
  // mesh: B[K][2][2], detJ[K], areaE0T[K][3], nuE0T[K][3][2]; neighbour
  // nb[K][3] (-1 none) with its local edge np[K][3]; markElemEdges[K][3]
  // (edge has a neighbour or a boundary id), markN[K][3] (Neumann),
  // markD[K][3] (Dirichlet edge)
  const double *B, *detJ, *areaE0T, *nuE0T;
  const int32_t *nb, *np;
  const uint8_t *markElemEdges, *markN, *markD;
  // point values (time-independent on the timed workloads): f, u1, u2 at the
  // element points [K][R2]; u1, u2, c_D at the edge points [K][3][R]
  const double *fq, *u1q, *u2q, *u1e, *u2e, *cDe;
  // limiter: globMTaylor and globMDiscTaylor blocks [K][N][N] (row i, col j),
  // phiTaylorV0T[K][3][N], vertex patches of local vertex n as CSR rows
  // ptr[n][K+1], idx[n][...], markV0TbdrD[K][3], Dirichlet vertex data cDV0T[K][3]
  const double *MT, *MDT, *phiTV0T;
  const int64_t* patchPtr[3];
  const int32_t* patchIdx[3];
  const uint8_t* markV0TbdrD;
  const double* cDV0T;
  const double* c0;  // initial state [K][N] (element-major, = reshape(cDisc', KN, 1))
};

struct refcpu_problem;
int refcpu_create(const refcpu_input* in, refcpu_problem** out);
int refcpu_run(refcpu_problem* h, int nsteps, double tau, int nthreads, double* seconds);
int refcpu_get_state(const refcpu_problem* h, double* c);
int refcpu_set_state(refcpu_problem* h, const double* c);
void refcpu_destroy(refcpu_problem* h);
}

namespace {

struct CSR {
  int n = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> col;
  std::vector<double> val;
};

// y = A x
void spmv(const CSR& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < A.n; ++r) {
    double s = 0.0;
    for (int64_t e = A.ptr[r]; e < A.ptr[r + 1]; ++e) s += A.val[e] * x[A.col[e]];
    y[r] = s;
  }
}

// C = sum_t w_t A_t (the sparse sum of Matlab's A + B): every row's column
// list is the merge of the terms' sorted lists, duplicates summed in term order
struct Term {
  double w;
  const CSR* A;
};
int merge_row(const Term* T, int nt, int r, int32_t* col, double* val) {
  const int32_t *c[4], *ce[4];
  const double* v[4];
  for (int t = 0; t < nt; ++t) {
    c[t] = T[t].A->col.data() + T[t].A->ptr[r];
    ce[t] = T[t].A->col.data() + T[t].A->ptr[r + 1];
    v[t] = T[t].A->val.data() + T[t].A->ptr[r];
  }
  int len = 0;
  for (;;) {
    int32_t m = std::numeric_limits<int32_t>::max();
    for (int t = 0; t < nt; ++t)
      if (c[t] < ce[t] && *c[t] < m) m = *c[t];
    if (m == std::numeric_limits<int32_t>::max()) break;
    if (col) {
      double s = 0.0;
      for (int t = 0; t < nt; ++t)
        if (c[t] < ce[t] && *c[t] == m) {
          s += T[t].w * *v[t]++;
          ++c[t];
        }
      col[len] = m;
      val[len] = s;
    } else {
      for (int t = 0; t < nt; ++t)
        if (c[t] < ce[t] && *c[t] == m) ++c[t];
    }
    ++len;
  }
  return len;
}

CSR sparse_sum(std::initializer_list<Term> terms) {
  const Term* T = terms.begin();
  const int nt = int(terms.size());
  const int n = T[0].A->n;
  CSR C;
  C.n = n;
  C.ptr.assign(size_t(n) + 1, 0);
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; ++r) C.ptr[r + 1] = merge_row(T, nt, r, nullptr, nullptr);
  for (int r = 0; r < n; ++r) C.ptr[r + 1] += C.ptr[r];
  C.col.resize(size_t(C.ptr[n]));
  C.val.resize(size_t(C.ptr[n]));
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; ++r) merge_row(T, nt, r, &C.col[size_t(C.ptr[r])], &C.val[size_t(C.ptr[r])]);
  return C;
}

// LU with partial pivoting of an n x n block (row-major, in place) and solve
struct LU {
  int n;
  double a[15 * 15];
  int piv[15];
  explicit LU(int n_) : n(n_) {}
  void factor(const double* A, double scale) {
    for (int i = 0; i < n * n; ++i) a[i] = scale * A[i];
    for (int k = 0; k < n; ++k) {
      int p = k;
      for (int i = k + 1; i < n; ++i)
        if (std::fabs(a[i * n + k]) > std::fabs(a[p * n + k])) p = i;
      piv[k] = p;
      if (p != k)
        for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[p * n + j]);
      const double d = a[k * n + k];
      for (int i = k + 1; i < n; ++i) {
        const double f = a[i * n + k] / d;
        a[i * n + k] = f;
        for (int j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
      }
    }
  }
  // P A = L U (the row interchanges swapped whole rows, L included): apply
  // all interchanges to x, then forward and back substitution
  void solve(double* x) const {
    for (int k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
    for (int k = 0; k < n; ++k)
      for (int i = k + 1; i < n; ++i) x[i] -= a[i * n + k] * x[k];
    for (int k = n - 1; k >= 0; --k) {
      for (int j = k + 1; j < n; ++j) x[k] -= a[k * n + j] * x[j];
      x[k] /= a[k * n + k];
    }
  }
};

}  // namespace

struct refcpu_problem {
  refcpu_input in;
  int K, N, R, R2;
  std::vector<double> c;  // current state [K][N]
  // per-stage data
  std::vector<double> fD, u1D, u2D, vn, KD, Lv;
  CSR G[2], Rm;

  double hatG(int i, int j, int l, int m) const { return in.hatG[((size_t(i) * N + j) * N + l) * 2 + m]; }

  // projectFuncCont2DataDisc: rhs = f(q) (W .* phi), x = rhs / hatM
  void project(const double* fq, std::vector<double>& out) const {
    out.assign(size_t(K) * N, 0.0);
    LU lu(N);
    // rhs / hatM == (hatM' \ rhs')': factor hatM' once for all rows
    std::vector<double> Mt(size_t(N) * N);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) Mt[i * N + j] = in.hatM[j * N + i];
    lu.factor(Mt.data(), 1.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k) {
      double* x = &out[size_t(k) * N];
      for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int q = 0; q < R2; ++q) s += fq[size_t(k) * R2 + q] * (in.W2[q] * in.phi2[q * N + i]);
        x[i] = s;
      }
      lu.solve(x);
    }
  }

  void preprocess_substep() {
    project(in.fq, fD);
    project(in.u1q, u1D);
    project(in.u2q, u2D);
    vn.resize(size_t(K) * 3 * R);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k)
      for (int n = 0; n < 3; ++n)
        for (int r = 0; r < R; ++r) {
          const size_t e = (size_t(k) * 3 + n) * R + r;
          vn[e] = in.nuE0T[(size_t(k) * 3 + n) * 2] * in.u1e[e] + in.nuE0T[(size_t(k) * 3 + n) * 2 + 1] * in.u2e[e];
        }
    // G1, G2: block (k, k) = sum_l u_m(k,l) [B(mo,mo) hatG(:,:,l,m) - B(mo,m) hatG(:,:,l,mo)]
    for (int m = 0; m < 2; ++m) {
      const int mo = 1 - m;
      const std::vector<double>& u = m == 0 ? u1D : u2D;
      CSR& A = G[m];
      A.n = K * N;
      A.ptr.resize(size_t(K) * N + 1);
      A.col.resize(size_t(K) * N * N);
      A.val.resize(size_t(K) * N * N);
#pragma omp parallel for schedule(static)
      for (int k = 0; k < K; ++k) {
        const double Bmm = in.B[size_t(k) * 4 + mo * 2 + mo], Bmo = in.B[size_t(k) * 4 + mo * 2 + m];
        for (int i = 0; i < N; ++i) {
          const size_t row = size_t(k) * N + i;
          A.ptr[row] = int64_t(row) * N;
          for (int j = 0; j < N; ++j) {
            double s = 0.0;
            for (int l = 0; l < N; ++l) {
              const double ul = u[size_t(k) * N + l];
              s = (s + (ul * Bmm) * hatG(i, j, l, m)) - (ul * Bmo) * hatG(i, j, l, mo);
            }
            A.col[row * N + j] = int32_t(size_t(k) * N + j);
            A.val[row * N + j] = s;
          }
        }
      }
      A.ptr[size_t(K) * N] = int64_t(K) * N * N;
    }
    // R: own block sum_{n,r} Rkn v [v > 0] hatRdiag(:,:,n,r); neighbour block
    // of edge n: sum_r Rkn v [v < 0] hatRoff(:,:,n,np,r); columns sorted
    Rm.n = K * N;
    Rm.ptr.assign(size_t(K) * N + 1, 0);
    std::vector<int> nblk(K);
    for (int k = 0; k < K; ++k) {
      int b = 1;
      for (int n = 0; n < 3; ++n) b += in.nb[size_t(k) * 3 + n] >= 0;
      nblk[k] = b;
    }
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < N; ++i) Rm.ptr[size_t(k) * N + i + 1] = nblk[k] * N;
    for (size_t r = 0; r < size_t(K) * N; ++r) Rm.ptr[r + 1] += Rm.ptr[r];
    Rm.col.resize(size_t(Rm.ptr[size_t(K) * N]));
    Rm.val.resize(Rm.col.size());
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k) {
      // blocks: own plus neighbours, by column
      int cols[4], which[4], nbk = 0;
      cols[nbk] = k;
      which[nbk++] = -1;
      for (int n = 0; n < 3; ++n)
        if (in.nb[size_t(k) * 3 + n] >= 0) {
          cols[nbk] = in.nb[size_t(k) * 3 + n];
          which[nbk++] = n;
        }
      for (int a = 1; a < nbk; ++a)
        for (int b = a; b > 0 && cols[b] < cols[b - 1]; --b) {
          std::swap(cols[b], cols[b - 1]);
          std::swap(which[b], which[b - 1]);
        }
      // upwind factors Rkn v [v > 0] (own block, edges with a neighbour or a
      // boundary id) and Rkn v [v < 0] (neighbour blocks)
      double dpos[3][8], dneg[3][8];
      for (int n = 0; n < 3; ++n) {
        const double Rkn = (in.markN[size_t(k) * 3 + n] ? 0.0 : 1.0) * in.areaE0T[size_t(k) * 3 + n];
        const bool me = in.markElemEdges[size_t(k) * 3 + n] != 0;
        for (int r = 0; r < R; ++r) {
          const double v = vn[(size_t(k) * 3 + n) * R + r];
          dpos[n][r] = Rkn * v * double(me && v > 0);
          dneg[n][r] = Rkn * v * double(v < 0);
        }
      }
      double blk[4][15 * 15];
      for (int b = 0; b < nbk; ++b) {
        double* o = blk[b];
        if (which[b] < 0) {
          for (int i = 0; i < N * N; ++i) o[i] = 0.0;
          for (int n = 0; n < 3; ++n)
            for (int r = 0; r < R; ++r) {
              const double d = dpos[n][r];
              for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) o[i * N + j] += d * in.hatRdiag[((size_t(i) * N + j) * 3 + n) * R + r];
            }
        } else {
          const int n = which[b], npn = in.np[size_t(k) * 3 + n];
          for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
              const double* h = &in.hatRoff[(((size_t(i) * N + j) * 3 + n) * 3 + npn) * R];
              double s = 0.0;
              for (int r = 0; r < R; ++r) s += dneg[n][r] * h[r];
              o[i * N + j] = s;
            }
        }
      }
      for (int i = 0; i < N; ++i) {
        int64_t o = Rm.ptr[size_t(k) * N + i];
        for (int b = 0; b < nbk; ++b)
          for (int j = 0; j < N; ++j, ++o) {
            Rm.col[size_t(o)] = int32_t(size_t(cols[b]) * N + j);
            Rm.val[size_t(o)] = blk[b][i * N + j];
          }
      }
    }
    // K_D and L = M f
    KD.assign(size_t(K) * N, 0.0);
    Lv.assign(size_t(K) * N, 0.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k) {
      for (int n = 0; n < 3; ++n) {
        const double Kkn = double(in.markD[size_t(k) * 3 + n]) * in.areaE0T[size_t(k) * 3 + n];
        for (int i = 0; i < N; ++i) {
          double s = 0.0;
          for (int r = 0; r < R; ++r) {
            const size_t e = (size_t(k) * 3 + n) * R + r;
            const double v = vn[e];
            s += (in.cDe[e] * (double(v < 0) * v)) * (in.W1[r] * in.phi1[(size_t(r) * N + i) * 3 + n]);
          }
          KD[size_t(k) * N + i] += Kkn * s;
        }
      }
      for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) s += in.detJ[k] * in.hatM[i * N + j] * fD[size_t(k) * N + j];
        Lv[size_t(k) * N + i] = s;
      }
    }
  }

  // globM \ y (every block detJ hatM factored, as a sparse direct solve does)
  void solve_M(std::vector<double>& y) const {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k) {
      LU lu(N);
      lu.factor(in.hatM, in.detJ[k]);
      lu.solve(&y[size_t(k) * N]);
    }
  }

  // y = blockdiag(detJ hatM) x
  void mul_M(const double* x, double* y) const {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) s += in.detJ[k] * in.hatM[i * N + j] * x[size_t(k) * N + j];
        y[size_t(k) * N + i] = s;
      }
  }

  // Disc -> Taylor: MDT \ (M x); Taylor -> Disc: M \ (MDT x)
  void disc_to_taylor(const double* x, double* t) const {
    mul_M(x, t);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k) {
      LU lu(N);
      lu.factor(&in.MDT[size_t(k) * N * N], 1.0);
      lu.solve(&t[size_t(k) * N]);
    }
  }
  void taylor_to_disc(const double* t, double* x) const {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k)
      for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int j = 0; j < N; ++j) s += in.MDT[(size_t(k) * N + i) * N + j] * t[size_t(k) * N + j];
        x[size_t(k) * N + i] = s;
      }
    std::vector<double> y(x, x + size_t(K) * N);
    solve_M(y);
    std::copy(y.begin(), y.end(), x);
  }

  // applySlopeLimiterTaylorLinear on T (in place), Dirichlet data dV (NaN: none)
  void limit_taylor_linear(double* T, const double* dV) const {
    std::vector<double> valC(K), valV(size_t(K) * 3);
    double vmin = std::numeric_limits<double>::infinity(), vmax = -vmin;
#pragma omp parallel for schedule(static) reduction(min : vmin) reduction(max : vmax)
    for (int k = 0; k < K; ++k) {
      valC[k] = T[size_t(k) * N];
      vmin = std::min(vmin, valC[k]);
      vmax = std::max(vmax, valC[k]);
      for (int n = 0; n < 3; ++n) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += T[size_t(k) * N + j] * in.phiTV0T[(size_t(k) * 3 + n) * N + j];
        valV[size_t(k) * 3 + n] = s;
      }
    }
    const double shiftPos = std::fabs(vmin) + 1, shiftNeg = std::fabs(vmax) + 1;
    const double tol = 1e-8;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k) {
      double alpha = std::numeric_limits<double>::infinity();
      for (int n = 0; n < 3; ++n) {
        double mn = std::numeric_limits<double>::infinity(), mx = -mn;
        for (int64_t e = in.patchPtr[n][k]; e < in.patchPtr[n][k + 1]; ++e) {
          const int32_t kk = in.patchIdx[n][e];
          mn = std::min(mn, valC[kk] - shiftNeg);
          mx = std::max(mx, valC[kk] + shiftPos);
        }
        mn = mn + shiftNeg;
        mx = mx - shiftPos;
        const double d0 = in.markV0TbdrD[size_t(k) * 3 + n] ? dV[size_t(k) * 3 + n] : std::nan("");
        mn = std::fmin(mn, d0);
        mx = std::fmax(mx, d0);
        const double d = valV[size_t(k) * 3 + n] - valC[k];
        const double dMin = mn - valC[k], dMax = mx - valC[k];
        double a = 1.0;
        if (d < dMin + tol) a = std::fmax(0.0, std::fmin(1.0, dMin / (d - tol)));
        if (d > dMax - tol) a = std::fmax(0.0, std::fmin(1.0, dMax / (d + tol)));
        alpha = std::min(alpha, a);
      }
      double* t = &T[size_t(k) * N];
      for (int j = 1; j < 3 && j < N; ++j) t[j] *= alpha;
      for (int j = 3; j < N; ++j) t[j] *= double(alpha == 1.0);
    }
  }

  void step(double tau) {
    const size_t KN = size_t(K) * N;
    std::vector<double> c0(c), acc, y(KN), cdot(KN), T(KN), TL(KN), nx(KN);
    static const double w3[3] = {0.0, 3.0 / 4, 1.0 / 3}, w2[2] = {0.0, 0.5};
    static const double A4[4] = {0.5, 0.5, 1.0, 1.0 / 6}, B4[3] = {1.0 / 6, 1.0 / 3, 1.0 / 3};
    const int S = in.ordRK;
    for (int s = 0; s < S; ++s) {
      preprocess_substep();  // time-independent point values: the same operators each sub-step
      // cdot = globM \ (L - KD - KN - sysA c), sysA = -G1 - G2 + R
      const CSR sysA = sparse_sum({Term{-1.0, &G[0]}, Term{-1.0, &G[1]}, Term{1.0, &Rm}});
      spmv(sysA, c.data(), y.data());
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < KN; ++e) cdot[e] = Lv[e] - KD[e] - y[e];
      solve_M(cdot);
      if (N > 3 && in.limiter) {
        // dc/dt limiting with selective mass lumping
        disc_to_taylor(cdot.data(), T.data());
        std::copy(T.begin(), T.end(), TL.begin());
        std::vector<double> nan3(size_t(K) * 3, std::nan(""));
        limit_taylor_linear(TL.data(), nan3.data());
#pragma omp parallel for schedule(static)
        for (int k = 0; k < K; ++k) {
          double dlt[15], out[15];
          for (int j = 0; j < N; ++j) dlt[j] = T[size_t(k) * N + j] - TL[size_t(k) * N + j];
          for (int i = 0; i < N; ++i) {
            double s = 0.0;
            const double rdii = 1.0 / in.MT[(size_t(k) * N + i) * N + i];  // diag(1 ./ diag(MT)) * MT
            for (int j = 0; j < N; ++j) s += (rdii * in.MT[(size_t(k) * N + i) * N + j]) * dlt[j];
            out[i] = TL[size_t(k) * N + i] + s;
          }
          for (int i = 0; i < N; ++i) T[size_t(k) * N + i] = out[i];
        }
        taylor_to_disc(T.data(), cdot.data());
      }
      if (S == 4) {
        if (s == 0) acc = c0;
        const double* base = s < 3 ? c0.data() : acc.data();
#pragma omp parallel for schedule(static)
        for (size_t e = 0; e < KN; ++e) nx[e] = base[e] + tau * A4[s] * cdot[e];
        if (s < 3) {
#pragma omp parallel for schedule(static)
          for (size_t e = 0; e < KN; ++e) acc[e] = acc[e] + tau * B4[s] * cdot[e];
        }
      } else {
        const double w = S == 3 ? w3[s] : S == 2 ? w2[s] : 0.0;
#pragma omp parallel for schedule(static)
        for (size_t e = 0; e < KN; ++e) nx[e] = w * c0[e] + (1 - w) * (c[e] + tau * cdot[e]);
      }
      if (in.limiter) {
        disc_to_taylor(nx.data(), T.data());
        limit_taylor_linear(T.data(), in.cDV0T);
        taylor_to_disc(T.data(), nx.data());
      }
      c.swap(nx);
    }
  }
};

extern "C" {

int refcpu_create(const refcpu_input* in, refcpu_problem** out) {
  if (!in || !out || in->K < 1 || in->N < 1 || in->N > 15 || in->R < 1 || in->R > 8 || in->ordRK < 1 || in->ordRK > 4) return 1;
  for (int e = 0; e < in->K * 3; ++e)
    if (in->markN[e]) return 2;  // Neumann data are not restated here
  refcpu_problem* h = new refcpu_problem;
  h->in = *in;
  h->K = in->K;
  h->N = in->N;
  h->R = in->R;
  h->R2 = in->R2;
  h->c.assign(in->c0, in->c0 + size_t(in->K) * in->N);
  *out = h;
  return 0;
}

int refcpu_run(refcpu_problem* h, int nsteps, double tau, int nthreads, double* seconds) {
  if (!h || nsteps < 0 || nthreads < 1) return 1;
  omp_set_num_threads(nthreads);
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < nsteps; ++s) h->step(tau);
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

int refcpu_get_state(const refcpu_problem* h, double* c) {
  if (!h || !c) return 1;
  std::copy(h->c.begin(), h->c.end(), c);
  return 0;
}

int refcpu_set_state(refcpu_problem* h, const double* c) {
  if (!h || !c) return 1;
  std::copy(c, c + h->c.size(), h->c.begin());
  return 0;
}

void refcpu_destroy(refcpu_problem* h) { delete h; }
}
