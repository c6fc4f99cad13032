// gfx950 kernels of the explicit DG advection step on triangles (FP64).
//
// One thread owns one element k.  Coefficients are structure-of-arrays
// c[i*K + k] (the column-major K x N layout of the reference's cDisc), so
// every per-element load and store of a wavefront is a contiguous 512-byte
// segment.  Reference tables are read through uniform (scalar) loads from
// one device buffer; the neighbour-trace table, whose index depends on the
// neighbour's local edge, is staged in LDS.
//
// Kernels:
//   stage_kernel       element term (core/assembleMatElemDphiPhiFuncDisc.m)
//                      + upwind edge term (core/assembleMatEdgePhiPhiValUpwind.m)
//                      + source + block mass solve + RK combination
//                      (advection/solveSubStep.m:62-83), optionally followed by
//                      the change to the Taylor basis of the limiter
//   vertex_minmax      per-vertex patch min / max of centroid values with the
//                      reference's global shift (core/computeMinMaxV0TElementPatch.m)
//   limit_kernel       alpha (core/computeVertexBasedLimiter.m), limited Taylor
//                      coefficients (core/applySlopeLimiterTaylorLinear.m), the
//                      dc/dt mass-lumping correction (advection/solveSubStep.m:75-78),
//                      change back to the DG basis, and for dc/dt the RK combination
//   to_taylor_kernel   DG -> Taylor for a stand-alone limiter call
//   centroid_kernel,   the limiter of the DG basis' linear part
//   limit_disc_linear  (core/applySlopeLimiterDiscLinear.m)
//   project / source / vnormal / l2 kernels for setup and checks.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace festung {

// The reference tables are read-only while any kernel runs.  Read through
// the constant address space, their uniform loads stay scalar loads whatever
// the kernel stores (or hides from the compiler) before them; through a plain
// pointer a store earlier on the path turns them into vector loads.
typedef __attribute__((address_space(4))) const double cdbl;
__device__ __forceinline__ cdbl* const_tab(const double* p) { return (cdbl*)p; }


constexpr int BLOCK = 256;

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Remap so that XCD x runs a contiguous range of logical blocks, in
// order: with elements stored along a space-filling curve, an XCD then works
// on a compact part of the mesh and neighbour / vertex-patch gathers across
// block boundaries hit its own L2.
constexpr int kXCDs = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int full = nb - nb % kXCDs;
  return b < full ? (b % kXCDs) * (full / kXCDs) + b / kXCDs : b;
}
__device__ __forceinline__ int logical_block() { return xcd_block(blockIdx.x, gridDim.x); }
__device__ __forceinline__ int gtid() { return logical_block() * int(blockDim.x) + threadIdx.x; }
// The same for a kernel launched with bs threads per workgroup over K
// elements (ceil(K / bs) workgroups), from the kernel's own arguments:
// blockDim.x and gridDim.x come from the dispatch packet through a 16-bit
// vector load, and the kernel's first memory wait then also waits for every
// load issued before it (the stage kernel's table loads, for one).
__device__ __forceinline__ int elem_blocks(int K, int bs) { return (K + bs - 1) / bs; }
__device__ __forceinline__ int elem_index(int K, int bs) {
  return xcd_block(blockIdx.x, elem_blocks(K, bs)) * bs + threadIdx.x;
}

// ---------------------------------------------------------------- Taylor --
// E_D(A): the degree-D symmetric power of a 2x2 matrix A acting on
// derivative multi-indices.  Row b (target derivative d1^(D-b) d2^b) holds
// the coefficients of (A11 X + A21)^(D-b) (A12 X + A22)^b, column b' the
// coefficient of X^(D-b') (source derivative d1^(D-b') d2^b').
// Row b of E_D(A): pw[k] = coefficient of X^k in (A11 X + A21)^(D-b)
// (A12 X + A22)^b, built one factor at a time; after t factors only
// pw[0 .. t] are non-zero, so only those are updated (no products with the
// zero padding, which IEEE arithmetic cannot drop).
template <int D>
__device__ __forceinline__ void sym_row(int b, double A11, double A21, double A12, double A22, double (&pw)[D + 1]) {
  pw[0] = 1.0;
#pragma unroll
  for (int t = 0; t < D; ++t) {
    const double c0 = t < D - b ? A21 : A22;
    const double c1 = t < D - b ? A11 : A12;
    pw[t + 1] = pw[t] * c1;
#pragma unroll
    for (int i = t; i >= 1; --i) pw[i] = pw[i] * c0 + pw[i - 1] * c1;
    pw[0] = pw[0] * c0;
  }
}

// y[0 .. D] = E_D(A) x[0 .. D], each row formed and contracted in turn
template <int D>
__device__ __forceinline__ void apply_sym(double A11, double A21, double A12, double A22, const double* x, double* y) {
#pragma unroll
  for (int b = 0; b <= D; ++b) {
    double pw[D + 1];
    sym_row<D>(b, A11, A21, A12, A22, pw);
    double s = 0.0;
#pragma unroll
    for (int bp = 0; bp <= D; ++bp) s += pw[D - bp] * x[bp];
    y[b] = s;
  }
}

template <int P>
struct Geo {
  double B11, B21, B12, B22, det, dX;
  __device__ __forceinline__ void load(const double* __restrict__ Bp, int k, int K) {
    B11 = Bp[k];
    B21 = Bp[k + K];
    B12 = Bp[k + 2 * K];
    B22 = Bp[k + 3 * K];
    det = B11 * B22 - B12 * B21;  // = 2|T| for a counter-clockwise element
    dX = 2.0 / sqrt(det);         // core/phiTaylorPhy.m:80
  }
};

// Indexing of the loop-ordered tables (TabOff::elemQ, edgeQ).
template <int P>
struct QIdx {
  static constexpr int N = Dims<P>::N, QB = 3 * N + 1, EB = N + 1;
  static __device__ __forceinline__ int w(int q) { return q * QB; }
  static __device__ __forceinline__ int phi(int q, int j) { return q * QB + 1 + j; }
  static __device__ __forceinline__ int d1(int q, int j) { return q * QB + 1 + N + j; }
  static __device__ __forceinline__ int d2(int q, int j) { return q * QB + 1 + 2 * N + j; }
  static __device__ __forceinline__ int ew(int n, int r) { return (n * Dims<P>::Re + r) * EB; }
  static __device__ __forceinline__ int ephi(int n, int r, int j) { return (n * Dims<P>::Re + r) * EB + 1 + j; }
};

// DG -> Taylor coefficients (exact change of basis between the orthonormal
// basis and the element's centroid-scaled, mean-free Taylor basis of
// core/phiTaylorPhy.m: T_0 = mean, T_(a,b) = d_xi^a d_eta^b u at the centroid,
// xi = dX (x - x_c)).  F maps DG coefficients to [mean; reference-coordinate
// derivatives at the centroid].
template <int P, class TabPtr>
__device__ __forceinline__ void dg_to_taylor(const double (&x)[Dims<P>::N], const Geo<P>& g, TabPtr F,
                                             double (&T)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  double y[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (taylor_nz(false, P, r, j)) s += F[r * N + j] * x[j];
    y[r] = s;
  }
  const double s = 1.0 / (g.det * g.dX);
  const double C11 = g.B22 * s, C12 = -g.B12 * s, C21 = -g.B21 * s, C22 = g.B11 * s;
  T[0] = y[0];
  if constexpr (P >= 1) apply_sym<1>(C11, C21, C12, C22, y + 1, T + 1);
  if constexpr (P >= 2) apply_sym<2>(C11, C21, C12, C22, y + 3, T + 3);
  if constexpr (P >= 3) apply_sym<3>(C11, C21, C12, C22, y + 6, T + 6);
  if constexpr (P >= 4) apply_sym<4>(C11, C21, C12, C22, y + 10, T + 10);
}

template <int P, class TabPtr>
__device__ __forceinline__ void taylor_to_dg(const double (&T)[Dims<P>::N], const Geo<P>& g, TabPtr Finv,
                                             double (&x)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  double z[N];
  const double A11 = g.dX * g.B11, A21 = g.dX * g.B21, A12 = g.dX * g.B12, A22 = g.dX * g.B22;
  z[0] = T[0];
  if constexpr (P >= 1) apply_sym<1>(A11, A21, A12, A22, T + 1, z + 1);
  if constexpr (P >= 2) apply_sym<2>(A11, A21, A12, A22, T + 3, z + 3);
  if constexpr (P >= 3) apply_sym<3>(A11, A21, A12, A22, T + 6, z + 6);
  if constexpr (P >= 4) apply_sym<4>(A11, A21, A12, A22, T + 10, z + 10);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (taylor_nz(true, P, i, j)) s += Finv[i * N + j] * z[j];
    x[i] = s;
  }
}

// The Taylor coefficients of degree <= 1 (mean and gradient) of dg_to_taylor,
// with the same operations in the same order.
template <int P, class TabPtr>
__device__ __forceinline__ void dg_to_taylor_lin(const double (&x)[Dims<P>::N], const Geo<P>& g, TabPtr F,
                                                 double (&T)[3]) {
  constexpr int N = Dims<P>::N;
  double y[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (taylor_nz(false, P, r, j)) s += F[r * N + j] * x[j];
    y[r] = s;
  }
  const double s = 1.0 / (g.det * g.dX);
  const double C11 = g.B22 * s, C12 = -g.B12 * s, C21 = -g.B21 * s, C22 = g.B11 * s;
  T[0] = y[0];
  apply_sym<1>(C11, C21, C12, C22, y + 1, T + 1);
}

// taylor_to_dg of a Taylor vector whose coefficients of degree >= 2 are zero
// (the 'linear' limiter's output where it limits): the products with the
// zero coefficients are left out, which leaves every sum unchanged (adding
// c * 0 to a sum does not change it).
template <int P, class TabPtr>
__device__ __forceinline__ void taylor_lin_to_dg(const double (&T)[3], const Geo<P>& g, TabPtr Finv,
                                                 double (&x)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  double z[3];
  const double A11 = g.dX * g.B11, A21 = g.dX * g.B21, A12 = g.dX * g.B12, A22 = g.dX * g.B22;
  z[0] = T[0];
  apply_sym<1>(A11, A21, A12, A22, T + 1, z + 1);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (taylor_nz(true, P, i, j)) s += Finv[i * N + j] * z[j];
    x[i] = s;
  }
}

// The 'linear' limiter without geometry: the first-order Taylor polynomial
// at the centroid, mean + grad f(x_c) . (x - x_c), is invariant under the
// affine map of the element, so (i) its deviation at vertex n is
// ghat . (vhat_n - chat), ghat the reference gradient at the reference
// centroid chat = (1/3, 1/3), and (ii) its DG coefficients are Finv's first
// three columns applied to (mean, ghat).  dg_to_taylor_lin / taylor_lin_to_dg
// reach the same values through B^-1 / dX and dX B, which cancel.
// ref_lin: (mean, ghat) = rows 0 .. 2 of F applied to the DG coefficients.
template <int P, class TabPtr>
__device__ __forceinline__ void ref_lin(const double (&x)[Dims<P>::N], TabPtr F, double (&y)[3]) {
  constexpr int N = Dims<P>::N;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (taylor_nz(false, P, r, j)) s += F[r * N + j] * x[j];
    y[r] = s;
  }
}
// ghat . (vhat_n - chat) at the vertices (0, 0), (1, 0), (0, 1)
__device__ __forceinline__ double ref_lin_dev(const double (&y)[3], int n) {
  constexpr double third = 1.0 / 3.0;
  return n == 0 ? -(y[1] + y[2]) * third : n == 1 ? (2.0 * y[1] - y[2]) * third : (2.0 * y[2] - y[1]) * third;
}
template <int P, class TabPtr>
__device__ __forceinline__ void ref_lin_to_dg(const double (&z)[3], TabPtr Finv, double (&x)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (taylor_nz(true, P, i, j)) s += Finv[i * N + j] * z[j];
    x[i] = s;
  }
}

// Element moments mu_(a,b) = mean over the element of xi^a eta^b, xi = dX
// (x - x_c), degree by degree.  For the linear form z = s xi + t eta the mean
// of z^D over a triangle is 2 D! / (D + 2)! h_D(z_0, z_1, z_2), with z_m its
// values at the vertices and h_D the complete homogeneous symmetric
// polynomial of degree D; hence mu_(D-b,b) = 2 / ((D+1)(D+2)) [s^(D-b) t^b]
// h_D / C(D, b).  The vertices are centred (e_1 = 0), so Newton's identities
// give h_D = -e_2 h_(D-2) + e_3 h_(D-3) (forms in (s, t); e_2, e_3 the
// elementary symmetric polynomials of the z_m, quadratic and cubic forms).
// Coefficient k of a form of degree d multiplies s^(d-k) t^k.
template <int D>
struct Deg {
  static constexpr int value = D;
};
__host__ __device__ constexpr double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

struct TriForms {
  double e2[3], e3[4];
};

template <int P>
__device__ __forceinline__ TriForms tri_forms(const Geo<P>& g) {
  // centred, scaled vertex coordinates (vertex 0 at -(b1 + b2) / 3)
  const double ex = (g.B11 + g.B12) / 3.0, ey = (g.B21 + g.B22) / 3.0;
  const double X[3] = {-ex * g.dX, (g.B11 - ex) * g.dX, (g.B12 - ex) * g.dX};
  const double Y[3] = {-ey * g.dX, (g.B21 - ey) * g.dX, (g.B22 - ey) * g.dX};
  TriForms f;
  f.e2[0] = X[0] * X[1] + X[0] * X[2] + X[1] * X[2];
  f.e2[1] = X[0] * Y[1] + Y[0] * X[1] + X[0] * Y[2] + Y[0] * X[2] + X[1] * Y[2] + Y[1] * X[2];
  f.e2[2] = Y[0] * Y[1] + Y[0] * Y[2] + Y[1] * Y[2];
  const double q0 = X[0] * X[1], q1 = X[0] * Y[1] + Y[0] * X[1], q2 = Y[0] * Y[1];  // z_0 z_1
  f.e3[0] = q0 * X[2];
  f.e3[1] = q0 * Y[2] + q1 * X[2];
  f.e3[2] = q1 * Y[2] + q2 * X[2];
  f.e3[3] = q2 * Y[2];
  return f;
}

// h (degree D) = -e2 hm2 + e3 hm3 (degrees D - 2, D - 3)
template <int D>
__device__ __forceinline__ void h_next(const TriForms& f, const double* hm2, const double* hm3, double* h) {
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (k - u >= 0 && k - u <= D - 2) s -= f.e2[u] * hm2[k - u];
    if constexpr (D >= 3) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k - u >= 0 && k - u <= D - 3) s += f.e3[u] * hm3[k - u];
    }
    h[k] = s;
  }
}

template <int D, int DMAX, typename F>
__device__ __forceinline__ void moment_degrees_from(const TriForms& tf, const double* hm3, const double* hm2,
                                                    const double* hm1, F& fn) {
  double h[D + 1], m[D + 1];
  h_next<D>(tf, hm2, hm3, h);
#pragma unroll
  for (int b = 0; b <= D; ++b) m[b] = h[b] * (2.0 / ((D + 1) * (D + 2)) / binom(D, b));
  fn(Deg<D>{}, m);
  if constexpr (D < DMAX) moment_degrees_from<D + 1, DMAX>(tf, hm2, hm1, h, fn);
}

// fn(Deg<D>, m) for D = 0 .. DMAX, m[b] = mu_(D-b,b)
template <int DMAX, int P, typename F>
__device__ __forceinline__ void moment_degrees(const Geo<P>& g, F&& fn) {
  const TriForms tf = tri_forms<P>(g);
  const double h0[1] = {1.0}, h1[2] = {0.0, 0.0};
  fn(Deg<0>{}, h0);
  if constexpr (DMAX >= 1) fn(Deg<1>{}, h1);
  if constexpr (DMAX >= 2) moment_degrees_from<2, DMAX>(tf, nullptr, h0, h1, fn);
}

// Taylor index i -> total degree d, exponent b of the second variable, and
// a! b! (a = d - b), for i = 0 .. 14 in the order (0,0),(1,0),(0,1),(2,0),...
__host__ __device__ constexpr int tdeg(int i) { return i == 0 ? 0 : i < 3 ? 1 : i < 6 ? 2 : i < 10 ? 3 : 4; }
__host__ __device__ constexpr int tb(int i) { return i - tdeg(i) * (tdeg(i) + 1) / 2; }
__host__ __device__ constexpr double fact4(int n) { return n == 4 ? 24.0 : n == 3 ? 6.0 : n == 2 ? 2.0 : 1.0; }
__host__ __device__ constexpr double tfab(int i) { return fact4(tdeg(i) - tb(i)) * fact4(tb(i)); }

// ------------------------------------------------------ min / max reduce --
// Workgroup min / max of the first s.F of the N Taylor coefficients Tc of
// each active thread -> partials (see MinMaxSink).  Plain stores: the
// consumer is a later kernel on the same stream.  Every thread of the
// workgroup must call this (it synchronises the workgroup).
// (bs: threads per workgroup, nb: workgroups, as launched)
template <int P>
__device__ __forceinline__ void sink_taylor(const double (&Tc)[Dims<P>::N], bool active, const MinMaxSink& s, int bs,
                                            int nb) {
  constexpr int FM = P * (P + 1) / 2 > 1 ? P * (P + 1) / 2 : 1;  // most fields a limiter type needs
  __shared__ double smn[FM][BLOCK / 64], smx[FM][BLOCK / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int f = 0; f < FM; ++f) {
    if (f < s.F) {
      double mn = active ? Tc[f] : INFINITY, mx = active ? Tc[f] : -INFINITY;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, off));
        mx = fmax(mx, __shfl_xor(mx, off));
      }
      if (lane == 0) {
        smn[f][w] = mn;
        smx[f][w] = mx;
      }
    }
  }
  __syncthreads();
  const int f = threadIdx.x;
  if (f < s.F) {
    double a = smn[f][0], b = smx[f][0];
    for (int i = 1; i < bs / 64; ++i) {
      a = fmin(a, smn[f][i]);
      b = fmax(b, smx[f][i]);
    }
    const size_t o = 2 * (size_t(f) * nb + blockIdx.x);
    s.partials[o] = a;
    s.partials[o + 1] = b;
  }
}

// ------------------------------------------------------------------- RK --
// out = a0 c0 + a1 c + b tau cdot + g acc;  accOut = d0 c0 + d1 acc + e tau cdot
// (rk_combine_pre below does the arithmetic; rk_combine first loads c0 and
// the accumulator where they are used -- all of them before the first store:
// accOut may be the accumulator itself, and loads placed after a store to it
// would each wait alone).  c is read only where a1 != 0 (classic RK4 has
// a1 = 0 at every stage, and its callers then leave c unloaded).
template <int N>
__device__ __forceinline__ void rk_combine_pre(const RKCoef& rk, double tau, const double (&c)[N],
                                               const double (&cdot)[N], const double (&c0v)[N],
                                               const double (&accv)[N], double* __restrict__ accOut, int K, int kk,
                                               bool active, double (&nx)[N]);
template <int N>
__device__ __forceinline__ void rk_combine(const RKCoef& rk, double tau, const double (&c)[N], const double (&cdot)[N],
                                           const double* __restrict__ c0, const double* __restrict__ accIn,
                                           double* __restrict__ accOut, int K, int kk, bool active, double (&nx)[N]) {
  double c0v[N], accv[N];
  if (rk.a0 != 0.0 || (accOut && rk.d0 != 0.0)) {
#pragma unroll
    for (int i = 0; i < N; ++i) c0v[i] = c0[i * K + kk];
  }
  if (rk.g != 0.0 || (accOut && rk.d1 != 0.0)) {
#pragma unroll
    for (int i = 0; i < N; ++i) accv[i] = accIn[i * K + kk];
  }
  rk_combine_pre<N>(rk, tau, c, cdot, c0v, accv, accOut, K, kk, active, nx);
}

// rk_combine with c0 and the accumulator already loaded (c0v, accv: read
// only where rk_combine reads c0 / accIn); the same operations
template <int N>
__device__ __forceinline__ void rk_combine_pre(const RKCoef& rk, double tau, const double (&c)[N],
                                               const double (&cdot)[N], const double (&c0v)[N],
                                               const double (&accv)[N], double* __restrict__ accOut, int K, int kk,
                                               bool active, double (&nx)[N]) {
  if (rk.a1 != 0.0) {
#pragma unroll
    for (int i = 0; i < N; ++i) nx[i] = rk.a1 * c[i] + rk.b * tau * cdot[i];
  } else {  // classic RK4: c is not read
#pragma unroll
    for (int i = 0; i < N; ++i) nx[i] = rk.b * tau * cdot[i];
  }
  if (rk.a0 != 0.0) {
#pragma unroll
    for (int i = 0; i < N; ++i) nx[i] += rk.a0 * c0v[i];
  }
  if (rk.g != 0.0) {
#pragma unroll
    for (int i = 0; i < N; ++i) nx[i] += rk.g * accv[i];
  }
  if (accOut && active) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double v = rk.e * tau * cdot[i];
      if (rk.d0 != 0.0) v += rk.d0 * c0v[i];
      if (rk.d1 != 0.0) v += rk.d1 * accv[i];
      accOut[i * K + kk] = v;
    }
  }
}

// Block mass solve  out = (det hatM)^-1 rhs  (det = 2|T|).  The basis is
// orthonormal and quadRule2D(max(2p, 1)) integrates phi_i phi_j exactly, so
// hatM is the identity up to the rounding of the rule: at p = 4, whose rule
// (core/triquad.m) is computed to full precision, within 6.4e-15 (summation
// rounding; the host checks it), and the product with hatM^-1 is skipped
// (225 products per element whose table streamed through the scalar cache);
// at p <= 3 the reference's 15-digit rule constants leave hatM off the
// identity by up to 1.7e-14, and the product stays.
template <int P, class TabPtr>
__device__ __forceinline__ void mass_solve(TabPtr minvR, const double (&rhs)[Dims<P>::N],
                                           double idet, double (&out)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  if constexpr (P == 4) {
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = rhs[i] * idet;
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += minvR[i * N + j] * rhs[j];
      out[i] = s * idet;
    }
  }
}

// ----------------------------------------------------------------- stage --

// Large element tables (p = 4) are staged in LDS: the scalar cache cannot
// hold them, and every wave streams the whole table.
template <int P>
struct ElemTabInLDS {
  static constexpr bool value = P >= 4;
  static constexpr int size = value ? Dims<P>::RE * (3 * Dims<P>::N + 1) : 1;
};

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Copy n doubles of a table into LDS with the whole workgroup; up to 8 loads
// per thread are in flight before the first store (a plain strided loop
// waits for each load in turn).
__device__ __forceinline__ void fill_lds(double* __restrict__ dst, const double* __restrict__ src, int n, int step) {
  constexpr int U = 8;
#pragma unroll 1
  for (int i0 = threadIdx.x; i0 < n; i0 += U * step) {
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = i0 + u * step < n ? src[i0 + u * step] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * step < n) dst[i0 + u * step] = v[u];
  }
}

// Element term  (G1 + G2) c  of the 64 elements e0 .. e0+63 of one wave on
// the matrix cores, as two GEMMs per 16-element group and 16-point block of
// the element term's rule (28 points at p = 4: the products are polynomials
// of degree 11, which that rule integrates exactly, as the reference's
// 64-point quadRule2D(8) does):
//   V[q][e] = sum_j phi_j(q) X_j[e]            for X = c, u1, u2
//   acc[i][e] += sum_q w_q d1phi_i(q) F1[q][e] + w_q d2phi_i(q) F2[q][e]
// with F1 = c (B22 u1 - B12 u2), F2 = c (B11 u2 - B21 u1) pointwise.  In the
// 16x16x4 FP64 MFMA, lane l holds A[l % 16][l / 16], B[l / 16][l % 16] and
// D[l / 16 + 4 t][l % 16] in accumulator register t, so register t of V is
// the B operand of the second GEMM for the points q = 16 qb + 4 t + l / 16
// (the table D1 / D2 is laid out to match): no data moves between lanes.
// The last block's rows past the rule are zero in A; its registers t with
// 16 qb + 4 t >= RG are skipped.
// The result goes to the wave's LDS buffer X as X[i][element - e0] (16 rows),
// where the one-element-per-lane code picks it up.  Every lane of the wave
// must call this (the MFMA is a wave operation).
template <int P>
__device__ __forceinline__ void elem_term_mfma(const StageArgs& a, double* __restrict__ X, int e0) {
  using M = MfmaElem<P>;
  constexpr int N = Dims<P>::N, S = M::S;
  const int K = a.m.K, l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
  // this lane's table operands (22 values) in registers, read once per wave
  // from the L2-resident table: no workgroup copy to LDS and no barrier
  // before the first MFMA
  const double* __restrict__ tm = a.tab + a.off.mfmaE;
  double tA[M::QB][S], tD1[M::NT], tD2[M::NT];
#pragma unroll
  for (int qb = 0; qb < M::QB; ++qb)
#pragma unroll
    for (int s = 0; s < S; ++s) tA[qb][s] = tm[(qb * S + s) * 64 + l];
#pragma unroll
  for (int t = 0; t < M::NT; ++t) {
    tD1[t] = tm[M::offD1 + t * 64 + l];
    tD2[t] = tm[M::offD2 + t * 64 + l];
  }
  // operands of group g in buffer g & 1; group g + 1's loads are issued
  // before group g's MFMAs
  double cB[2][S], u1B[2][S], u2B[2][S], Bg[2][4];
  auto load = [&](int g, int b) {
    const int e = e0 + 16 * g + lr < K ? e0 + 16 * g + lr : 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int j = 4 * s + lk;
      const bool in = j < N;
      cB[b][s] = in ? a.c[j * K + e] : 0.0;
      u1B[b][s] = in ? a.u1[j * K + e] : 0.0;
      u2B[b][s] = in ? a.u2[j * K + e] : 0.0;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) Bg[b][m] = a.m.B[e + m * K];
  };
  load(0, 0);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int b = g & 1;
    if (g < 3) load(g + 1, b ^ 1);
    const double B11 = Bg[b][0], B21 = Bg[b][1], B12 = Bg[b][2], B22 = Bg[b][3];
    dbl4 r1 = {0.0, 0.0, 0.0, 0.0}, r2 = r1;
#pragma unroll
    for (int qb = 0; qb < M::QB; ++qb) {
      dbl4 vc = {0.0, 0.0, 0.0, 0.0}, v1 = vc, v2 = vc;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const double ta = tA[qb][s];
        vc = __builtin_amdgcn_mfma_f64_16x16x4f64(ta, cB[b][s], vc, 0, 0, 0);
        v1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ta, u1B[b][s], v1, 0, 0, 0);
        v2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ta, u2B[b][s], v2, 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (t >= M::steps(qb)) break;
        const double f1 = vc[t] * (B22 * v1[t] - B12 * v2[t]);
        const double f2 = vc[t] * (B11 * v2[t] - B21 * v1[t]);
        r1 = __builtin_amdgcn_mfma_f64_16x16x4f64(tD1[qb * 4 + t], f1, r1, 0, 0, 0);
        r2 = __builtin_amdgcn_mfma_f64_16x16x4f64(tD2[qb * 4 + t], f2, r2, 0, 0, 0);
      }
    }
    const dbl4 r = r1 + r2;
#pragma unroll
    for (int t = 0; t < 4; ++t) X[(lk + 4 * t) * 64 + 16 * g + lr] = r[t];
  }
}

template <int P, int MODE>
__global__ __launch_bounds__(BLOCK) void stage_kernel(StageArgs a) {
  constexpr int N = Dims<P>::N, RE = Dims<P>::RE, Re = Dims<P>::Re;
  constexpr bool MF = MfmaElem<P>::value;
  __shared__ double sTheta[Re * N * 9];
  __shared__ double sQ[MF ? 1 : ElemTabInLDS<P>::size];
  extern __shared__ double sXchg[];  // MF: MfmaElem<P>::xchg per wave (dynamic)
  // Only the edge term reads LDS (the neighbour-trace table; at p = 4 the
  // element term runs on the matrix cores with its table operands in
  // registers): the table's loads are issued first, the element's own loads
  // (and the matrix-core element term) behind them, and the table is written
  // to LDS after those, with a barrier that waits for LDS only -- so the
  // element's loads stay in flight across it instead of starting after the
  // workgroup's table fill.
  constexpr bool kLateFill = MF || !ElemTabInLDS<P>::value;
  constexpr int NTH = Re * N * 9, TPT = (NTH + 63) / 64;
  [[maybe_unused]] double thv[kLateFill ? TPT : 1];
  if constexpr (kLateFill) {
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const int i = int(threadIdx.x) + u * a.m.bs;
      thv[u] = i < NTH ? a.tab[a.off.theta + i] : 0.0;
    }
  } else {
    fill_lds(sTheta, a.tab + a.off.theta, NTH, a.m.bs);
    if constexpr (ElemTabInLDS<P>::value) fill_lds(sQ, a.tab + a.off.elemQ, ElemTabInLDS<P>::size, a.m.bs);
    __syncthreads();
  }
  const double* __restrict__ EQ = ElemTabInLDS<P>::value ? sQ : a.tab + a.off.elemQ;

  const int K = a.m.K;
  const int k = elem_index(K, a.m.bs);
  const bool active = k < K;
  const int kk = active ? k : 0;
  const double* __restrict__ T = a.tab;

  // element term  (G1 + G2) c  in quadrature form on the element term's
  // rule: on the matrix cores, into the wave's LDS buffer, before the edge
  // term; else in this lane's loop below.  Either way acc = edge term +
  // element term in the end.  (With 38 MFMAs per 16 elements, staggering the
  // two waves of a SIMD -- one forming the element term first, the other
  // last -- no longer pays: 32.7 against 34.4 us at K = 131 072.)
  double c[N], acc[N];
  [[maybe_unused]] double* X = sXchg + (threadIdx.x >> 6) * MfmaElem<P>::xchg;
  if constexpr (MF) elem_term_mfma<P>(a, X, k - (threadIdx.x & 63));
#pragma unroll
  for (int i = 0; i < N; ++i) {
    c[i] = a.c[i * K + kk];
    acc[i] = 0.0;
  }
  // the DG outputs at p <= 3 need only 1 / det (the element term's geometry
  // is in the velocity weights, the edge term's in the flux factors)
  constexpr bool kRdetOnly = !MF && (MODE == OUT_DOT_DG || MODE == OUT_DOT_DG_CEN || MODE == OUT_RK_DG);
  [[maybe_unused]] Geo<P> g;
  double rdet;
  if constexpr (kRdetOnly) {
    rdet = a.m.rdet[kk];
  } else {
    g.load(a.m.B, kk, K);
    rdet = 1.0 / g.det;
  }
  // RK operands (c_0, the accumulator) and the element's time step with the
  // element's other loads: read where they are combined, each load came
  // after the previous coefficient's store (the output may alias them) and
  // waited alone.  (p = 4: the registers are not there.)
  constexpr bool kRKPre = (MODE == OUT_RK_DG || MODE == OUT_RK_TAYLOR) && P <= 3;
  [[maybe_unused]] double c0v[kRKPre ? N : 1], accv[kRKPre ? N : 1];
  [[maybe_unused]] double tkp = a.tau;
  if constexpr (kRKPre) {
    const RKCoef& rk = a.rk;
    if (rk.a0 != 0.0 || (a.accOut && rk.d0 != 0.0)) {
#pragma unroll
      for (int i = 0; i < N; ++i) c0v[i] = a.c0[i * K + kk];
    }
    if (rk.g != 0.0 || (a.accOut && rk.d1 != 0.0)) {
#pragma unroll
      for (int i = 0; i < N; ++i) accv[i] = a.acc[i * K + kk];
    }
    if (a.tauK) tkp = a.tauK[kk];
  }

  // Neighbour ids are requested up front.  At p <= 2 the neighbours'
  // coefficients are gathered one edge ahead (edge 0's while the element
  // term is formed, edge n + 1's while edge n is), so the dependent gathers
  // are in flight behind arithmetic; at p >= 3 the registers are not there
  // and each edge gathers its neighbour when it starts.  A missing neighbour
  // reads the element's own coefficients (in bounds; the select below
  // discards them).
  // (p = 3 keeps the gathers, and the id loads, per edge: measured faster)
  constexpr bool kAhead = P <= 2;
  constexpr bool kSelect = P != 3;
  const int fl = kSelect ? a.m.flags[kk] : 0;
  int nbs[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) nbs[n] = kSelect ? a.m.nbr[n * K + kk] : -1;
  double cnb[kAhead ? 2 : 1][N];
  auto gather = [&](int n, double (&dst)[N]) {
    const int nbk = nbs[n] >= 0 ? nbs[n] : kk;
#pragma unroll
    for (int j = 0; j < N; ++j) dst[j] = a.c[j * K + nbk];
  };
  if constexpr (kAhead) gather(0, cnb[0]);
  if constexpr (kLateFill) {
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
      const int i = int(threadIdx.x) + u * a.m.bs;
      if (i < NTH) sTheta[i] = thv[u];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }

  if constexpr (!MF && P >= 1) {
    // the velocity enters through its weighted factors at the points (uq)
    constexpr int UQ = RE <= 6 ? RE : 1;  // p >= 3: not unrolled (measured fastest at p = 3, 4)
#pragma unroll UQ
    for (int q = 0; q < RE; ++q) {
      double cq = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) cq += c[j] * EQ[QIdx<P>::phi(q, j)];
      const double f1 = cq * a.uq[q * K + kk];
      const double f2 = cq * a.uq[(RE + q) * K + kk];
#pragma unroll
      for (int i = 1; i < N; ++i) acc[i] += EQ[QIdx<P>::d1(q, i)] * f1 + EQ[QIdx<P>::d2(q, i)] * f2;
    }
  }

  // edge term  -R c  (upwind; interior neighbours gathered, no atomics).
  // Except at p = 3, both traces are formed at every point and the upwind
  // one selected: the upwind side differs from lane to lane, and branching on
  // it made every wave run both sides anyway, with the branch overhead on top
  // (p = 1, 2, 4: 20.3 -> 16.4, 37.3 -> 33.0, 49.1 -> 43.7 us per launch at
  // the benchmark sizes); at p = 3 the branches are faster (24.0 against
  // 26.0 us).  The same trace is selected either way: the same bits.
  auto edge = [&](int n, const double (&cn)[N]) {
    const int np = flag_np(fl, n);
    const bool outOn = flag_out(fl, n);
    const bool inOn = nbs[n] >= 0;
#pragma unroll
    for (int r = 0; r < Re; ++r) {
      const double v = a.vnw[(r * 3 + n) * K + kk];  // |E| w_r u.nu: the sign of u.nu
      const double* th = sTheta + r + n * Re * N + np * Re * N * 3;
      double tr = 0.0;
      if constexpr (kSelect) {
        double to = 0.0, tn = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) to += c[j] * T[a.off.edgeQ + QIdx<P>::ephi(n, r, j)];
#pragma unroll
        for (int j = 0; j < N; ++j) tn += cn[j] * th[j * Re];
        tr = (v > 0.0 && outOn) ? to : ((v < 0.0 && inOn) ? tn : 0.0);
      } else {
        if (v > 0.0 && outOn) {
#pragma unroll
          for (int j = 0; j < N; ++j) tr += c[j] * T[a.off.edgeQ + QIdx<P>::ephi(n, r, j)];
        } else if (v < 0.0 && inOn) {
#pragma unroll
          for (int j = 0; j < N; ++j) tr += cn[j] * th[j * Re];
        }
      }
      const double fx = v * tr;
#pragma unroll
      for (int i = 0; i < N; ++i) acc[i] -= T[a.off.edgeQ + QIdx<P>::ephi(n, r, i)] * fx;
    }
  };
  if constexpr (kAhead) {
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      if (n < 2) gather(n + 1, cnb[(n + 1) & 1]);
      edge(n, cnb[n & 1]);
    }
  } else if constexpr (kSelect) {
#pragma unroll 1
    for (int n = 0; n < 3; ++n) {
      gather(n, cnb[0]);
      edge(n, cnb[0]);
    }
  } else {
    // p = 3: the branches, the per-element flag and per-edge id loads here
    // and only real neighbours gathered, written out (this form schedules
    // best at p = 3)
    const int fl3 = a.m.flags[kk];
#pragma unroll 1
    for (int n = 0; n < 3; ++n) {
      const int nb = a.m.nbr[n * K + kk];
      const int np = flag_np(fl3, n);
      const bool outOn = flag_out(fl3, n);
      double cn[N];
      if (nb >= 0) {
#pragma unroll
        for (int j = 0; j < N; ++j) cn[j] = a.c[j * K + nb];
      }
#pragma unroll
      for (int r = 0; r < Re; ++r) {
        const double v = a.vnw[(r * 3 + n) * K + kk];
        double tr = 0.0;
        if (v > 0.0 && outOn) {
#pragma unroll
          for (int j = 0; j < N; ++j) tr += c[j] * T[a.off.edgeQ + QIdx<P>::ephi(n, r, j)];
        } else if (v < 0.0 && nb >= 0) {
          const double* th = sTheta + r + n * Re * N + np * Re * N * 3;
#pragma unroll
          for (int j = 0; j < N; ++j) tr += cn[j] * th[j * Re];
        }
        const double fx = v * tr;
#pragma unroll
        for (int i = 0; i < N; ++i) acc[i] -= T[a.off.edgeQ + QIdx<P>::ephi(n, r, i)] * fx;
      }
    }
  }
  if constexpr (MF) {
    __builtin_amdgcn_wave_barrier();  // X is written and read by this wave only
#pragma unroll
    for (int i = 1; i < N; ++i) acc[i] += X[i * 64 + (threadIdx.x & 63)];
  }
  if (a.rhs) {
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] += a.rhs[i * K + kk];
  }
  // block mass solve: cdot = (detJ hatM)^-1 rhs
  double cdot[N];
  mass_solve<P>(T + a.off.hatMinvR, acc, rdet, cdot);

  if constexpr (MODE == OUT_DOT_DG) {
    if (active)
#pragma unroll
      for (int i = 0; i < N; ++i) a.out[i * K + k] = cdot[i];
    return;
  } else if constexpr (MODE == OUT_DOT_DG_CEN) {
    // dc/dt in DG for limit_dg_kernel, with the min / max partials of its
    // Taylor mean (dg_to_taylor's T_0 = F_00 x_0; F_00 read before the
    // stores, which would turn its scalar load into a waited vector load)
    double Tc[N];
    Tc[0] = T[a.off.FR] * cdot[0];
    if (active)
#pragma unroll
      for (int i = 0; i < N; ++i) a.out[i * K + k] = cdot[i];
#pragma unroll
    for (int i = 1; i < N; ++i) Tc[i] = 0.0;
    sink_taylor<P>(Tc, active, a.sink, a.m.bs, elem_blocks(K, a.m.bs));
  } else if constexpr (MODE == OUT_DOT_TAYLOR) {
    double Tc[N];
    dg_to_taylor<P>(cdot, g, const_tab(T + a.off.FR), Tc);
    if (active) {
#pragma unroll
      for (int i = 0; i < N; ++i) a.out[i * K + k] = Tc[i];
    }
    sink_taylor<P>(Tc, active, a.sink, a.m.bs, elem_blocks(K, a.m.bs));
  } else {
    // RK combination (advection/solveSubStep.m:83 for the SSP weights)
    double nx[N];
    if constexpr (kRKPre) {
      rk_combine_pre<N>(a.rk, tkp, c, cdot, c0v, accv, a.accOut, K, kk, active, nx);
    } else {
      const double tk = a.tauK ? a.tauK[kk] : a.tau;
      rk_combine<N>(a.rk, tk, c, cdot, a.c0, a.acc, a.accOut, K, kk, active, nx);
    }
    if constexpr (MODE == OUT_RK_DG) {
      if (active)
#pragma unroll
        for (int i = 0; i < N; ++i) a.out[i * K + k] = nx[i];
    } else {
      double Tc[N];
      dg_to_taylor<P>(nx, g, const_tab(T + a.off.FR), Tc);
      if (active) {
#pragma unroll
        for (int i = 0; i < N; ++i) a.out[i * K + k] = Tc[i];
      }
      sink_taylor<P>(Tc, active, a.sink, a.m.bs, elem_blocks(K, a.m.bs));
    }
  }
}

// -------------------------------------------------------- vertex min/max --

__global__ __launch_bounds__(BLOCK) void vertex_minmax_kernel(VertexArgs a) {
  if (blockIdx.x == 0) {
    // global min / max of each coefficient from the producer's partials;
    // loads are issued in batches of 8 before they are reduced
    __shared__ double smn[BLOCK / 64], smx[BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int f = 0; f < a.F; ++f) {
      const double* part = a.partials + 2 * size_t(f) * a.nPartials;
      double mn = INFINITY, mx = -INFINITY;
      for (int i0 = 0; i0 < a.nPartials; i0 += 8 * BLOCK) {
        // (out-of-range slots read partial 0 and are masked after the
        // loads, so that the eight loads issue together instead of each
        // behind a branch and its own wait)
        double lo[8], hi[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int i = i0 + u * BLOCK + threadIdx.x;
          const int ic = i < a.nPartials ? i : 0;
          lo[u] = part[2 * ic];
          hi[u] = part[2 * ic + 1];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const bool ok = i0 + u * BLOCK + int(threadIdx.x) < a.nPartials;
          lo[u] = ok ? lo[u] : INFINITY;
          hi[u] = ok ? hi[u] : -INFINITY;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          mn = fmin(mn, lo[u]);
          mx = fmax(mx, hi[u]);
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, off));
        mx = fmax(mx, __shfl_xor(mx, off));
      }
      if (lane == 0) {
        smn[w] = mn;
        smx[w] = mx;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 1; i < BLOCK / 64; ++i) {
          mn = fmin(mn, smn[i]);
          mx = fmax(mx, smx[i]);
        }
        a.gmm[2 * f] = mn;
        a.gmm[2 * f + 1] = mx;
      }
      __syncthreads();  // smn / smx are reused for the next coefficient
    }
    return;
  }
  // workgroups 1 .. numV / BLOCK: one vertex per thread
  const int v = xcd_block(blockIdx.x - 1, elem_blocks(a.numV, BLOCK)) * BLOCK + threadIdx.x;
  if (v >= a.numV) return;
  for (int j0 = 0; j0 < a.maxDeg; j0 += 8) {
    int id[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) id[u] = j0 + u < a.maxDeg ? a.v2tPad[(j0 + u) * a.numV + v] : -1;
    for (int f = 0; f < a.F; ++f) {
      const double* fld = a.fields + size_t(f) * a.K;
      const double sc = f == 0 ? a.scale0 : 1.0;
      double m = j0 == 0 ? INFINITY : a.vmin[size_t(f) * a.numV + v];
      double M = j0 == 0 ? -INFINITY : a.vmax[size_t(f) * a.numV + v];
      // all gathers first (a padding slot reads element 0, masked below):
      // behind a per-slot branch each one waited for the one before
      double cv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) cv[u] = sc * fld[id[u] >= 0 ? id[u] : 0];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (id[u] >= 0) {
          m = fmin(m, cv[u]);
          M = fmax(M, cv[u]);
        }
      }
      a.vmin[size_t(f) * a.numV + v] = m;
      a.vmax[size_t(f) * a.numV + v] = M;
    }
  }
}

// ----------------------------------------------------------------- limit --

// Selective mass lumping of the dc/dt limiter (advection/solveSubStep.m:75-78):
// lim += diag(MT)^-1 MT (orig - lim) on the Taylor coefficients, MT the
// element's Taylor mass matrix.
template <int P>
__device__ __forceinline__ void lump_correct(const Geo<P>& g, const double (&x)[Dims<P>::N], double (&lim)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  // corr_i = f_i / (mu_2i - mu_i^2) * (sum_j mu_(i+j) d_j - mu_i sum_j mu_j d_j),
  // d_j = (orig_j - lim_j) / f_j, f = a! b!: row i of diag(MT)^-1 MT applied
  // to (orig - lim), with MT the element's Taylor mass matrix
  double dj[N];
  dj[0] = 0.0;
#pragma unroll
  for (int j = 1; j < N; ++j) dj[j] = (x[j] - lim[j]) * (1.0 / tfab(j));
  // one pass over the moment degrees D = 0 .. 2p: the moments of degree
  // <= p are kept (mlo); degree D adds mu_(i+j) d_j to the rows i with
  // deg i + deg j = D and sets den_i = mu_(2i) - mu_i^2 at D = 2 deg i
  double mlo[N], s[N], den[N];
#pragma unroll
  for (int i = 0; i < N; ++i) s[i] = den[i] = 0.0;
  moment_degrees<2 * P>(g, [&](auto dc, const double* m) {
    constexpr int D = decltype(dc)::value;
    if constexpr (D <= P) {
#pragma unroll
      for (int b = 0; b <= D; ++b) mlo[D * (D + 1) / 2 + b] = m[b];
    }
    if constexpr (D >= 2) {
#pragma unroll
      for (int i = 1; i < N; ++i) {
#pragma unroll
        for (int j = 1; j < N; ++j)
          if (tdeg(i) + tdeg(j) == D) s[i] += m[tb(i) + tb(j)] * dj[j];
        if (2 * tdeg(i) == D) den[i] = m[2 * tb(i)] - mlo[i] * mlo[i];
      }
    }
  });
  double S0 = 0.0;
#pragma unroll
  for (int j = 1; j < N; ++j) S0 += mlo[j] * dj[j];
#pragma unroll
  for (int i = 1; i < N; ++i) lim[i] += (s[i] - mlo[i] * S0) * tfab(i) / den[i];
}

// ------------------------------------------------ double-double (fix-up) --
// The dc/dt limiter's branch for limited elements takes every Taylor
// coefficient through the closed-form change of basis and back, which
// rounds like cond(B)^p eps, the rounding of the tables F and Finv
// included.  On badly shaped elements (the host lists those with cond(B)
// above a threshold; none on structured meshes) the same operations run in
// double-double arithmetic (about 106 bits: Dekker's and Knuth's error-free
// sum and product, Bailey's double-double; the products through FMA), with
// F and Finv as double-double tables, in limit_dot_fixup_kernel, which hands
// the limited dc/dt to the limiter.
struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd mkdd(double h) { return dd{h, 0.0}; }
__device__ __forceinline__ dd two_sum(double a, double b) {
  const double s = a + b, v = s - a;
  return dd{s, (a - (s - v)) + (b - v)};
}
__device__ __forceinline__ dd fast_two_sum(double a, double b) {
  const double s = a + b;
  return dd{s, b - (s - a)};
}
__device__ __forceinline__ dd operator+(const dd& a, const dd& b) {
  const dd s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
  double c = s.lo + t.hi;
  const dd v = fast_two_sum(s.hi, c);
  c = t.lo + v.lo;
  return fast_two_sum(v.hi, c);
}
__device__ __forceinline__ dd operator-(const dd& a) { return dd{-a.hi, -a.lo}; }
__device__ __forceinline__ dd operator-(const dd& a, const dd& b) { return a + (-b); }
__device__ __forceinline__ dd operator*(const dd& a, const dd& b) {
  const double p = a.hi * b.hi;
  double e = __builtin_fma(a.hi, b.hi, -p);
  e = __builtin_fma(a.hi, b.lo, e);
  e = __builtin_fma(a.lo, b.hi, e);
  return fast_two_sum(p, e);
}
__device__ __forceinline__ dd operator/(const dd& a, const dd& b) {
  const double q1 = a.hi / b.hi;
  dd r = a - b * mkdd(q1);
  const double q2 = r.hi / b.hi;
  r = r - b * mkdd(q2);
  const double q3 = r.hi / b.hi;
  return fast_two_sum(q1, q2) + mkdd(q3);
}
__device__ __forceinline__ dd dd_sqrt(const dd& a) {  // one Newton step from the double root
  const double x = sqrt(a.hi);
  return mkdd(x) + (a - mkdd(x) * mkdd(x)) / mkdd(2.0 * x);
}

template <int P>
struct GeoDD {
  double B11, B21, B12, B22;
  dd det, dX;
  __device__ __forceinline__ void load(const double* __restrict__ Bp, int k, int K) {
    B11 = Bp[k];
    B21 = Bp[k + K];
    B12 = Bp[k + 2 * K];
    B22 = Bp[k + 3 * K];
    det = mkdd(B11) * mkdd(B22) - mkdd(B12) * mkdd(B21);
    dX = mkdd(2.0) / dd_sqrt(det);
  }
};

template <int D>
__device__ __forceinline__ void apply_sym_dd(const dd& A11, const dd& A21, const dd& A12, const dd& A22, const dd* x,
                                             dd* y) {
#pragma unroll
  for (int b = 0; b <= D; ++b) {
    dd pw[D + 1];
    pw[0] = mkdd(1.0);
#pragma unroll
    for (int t = 0; t < D; ++t) {
      const dd c0 = t < D - b ? A21 : A22;
      const dd c1 = t < D - b ? A11 : A12;
      pw[t + 1] = pw[t] * c1;
#pragma unroll
      for (int i = t; i >= 1; --i) pw[i] = pw[i] * c0 + pw[i - 1] * c1;
      pw[0] = pw[0] * c0;
    }
    dd s = mkdd(0.0);
#pragma unroll
    for (int bp = 0; bp <= D; ++bp) s = s + pw[D - bp] * x[bp];
    y[b] = s;
  }
}

template <int P>
__device__ __forceinline__ void dg_to_taylor_dd(const double (&x)[Dims<P>::N], const GeoDD<P>& g,
                                                cdbl* F, cdbl* Flo,
                                                dd (&T)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  dd y[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    dd s = mkdd(0.0);
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (taylor_nz(false, P, r, j)) s = s + dd{F[r * N + j], Flo[r * N + j]} * mkdd(x[j]);
    y[r] = s;
  }
  const dd sc = mkdd(1.0) / (g.det * g.dX);
  const dd C11 = mkdd(g.B22) * sc, C12 = -(mkdd(g.B12) * sc), C21 = -(mkdd(g.B21) * sc), C22 = mkdd(g.B11) * sc;
  T[0] = y[0];
  if constexpr (P >= 1) apply_sym_dd<1>(C11, C21, C12, C22, y + 1, T + 1);
  if constexpr (P >= 2) apply_sym_dd<2>(C11, C21, C12, C22, y + 3, T + 3);
  if constexpr (P >= 3) apply_sym_dd<3>(C11, C21, C12, C22, y + 6, T + 6);
  if constexpr (P >= 4) apply_sym_dd<4>(C11, C21, C12, C22, y + 10, T + 10);
}

template <int P>
__device__ __forceinline__ void taylor_to_dg_dd(const dd (&T)[Dims<P>::N], const GeoDD<P>& g,
                                                cdbl* Finv, cdbl* Finvlo,
                                                dd (&x)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  dd z[N];
  const dd A11 = g.dX * mkdd(g.B11), A21 = g.dX * mkdd(g.B21), A12 = g.dX * mkdd(g.B12), A22 = g.dX * mkdd(g.B22);
  z[0] = T[0];
  if constexpr (P >= 1) apply_sym_dd<1>(A11, A21, A12, A22, T + 1, z + 1);
  if constexpr (P >= 2) apply_sym_dd<2>(A11, A21, A12, A22, T + 3, z + 3);
  if constexpr (P >= 3) apply_sym_dd<3>(A11, A21, A12, A22, T + 6, z + 6);
  if constexpr (P >= 4) apply_sym_dd<4>(A11, A21, A12, A22, T + 10, z + 10);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    dd s = mkdd(0.0);
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (taylor_nz(true, P, i, j)) s = s + dd{Finv[i * N + j], Finvlo[i * N + j]} * z[j];
    x[i] = s;
  }
}

// tri_forms / h_next / moment_degrees in double-double
struct TriFormsDD {
  dd e2[3], e3[4];
};
template <int P>
__device__ __forceinline__ TriFormsDD tri_forms_dd(const GeoDD<P>& g) {
  const dd three = mkdd(3.0);
  const dd ex = (mkdd(g.B11) + mkdd(g.B12)) / three, ey = (mkdd(g.B21) + mkdd(g.B22)) / three;
  const dd X[3] = {-(ex * g.dX), (mkdd(g.B11) - ex) * g.dX, (mkdd(g.B12) - ex) * g.dX};
  const dd Y[3] = {-(ey * g.dX), (mkdd(g.B21) - ey) * g.dX, (mkdd(g.B22) - ey) * g.dX};
  TriFormsDD f;
  f.e2[0] = X[0] * X[1] + X[0] * X[2] + X[1] * X[2];
  f.e2[1] = X[0] * Y[1] + Y[0] * X[1] + X[0] * Y[2] + Y[0] * X[2] + X[1] * Y[2] + Y[1] * X[2];
  f.e2[2] = Y[0] * Y[1] + Y[0] * Y[2] + Y[1] * Y[2];
  const dd q0 = X[0] * X[1], q1 = X[0] * Y[1] + Y[0] * X[1], q2 = Y[0] * Y[1];
  f.e3[0] = q0 * X[2];
  f.e3[1] = q0 * Y[2] + q1 * X[2];
  f.e3[2] = q1 * Y[2] + q2 * X[2];
  f.e3[3] = q2 * Y[2];
  return f;
}
template <int D>
__device__ __forceinline__ void h_next_dd(const TriFormsDD& f, const dd* hm2, const dd* hm3, dd* h) {
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    dd s = mkdd(0.0);
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (k - u >= 0 && k - u <= D - 2) s = s - f.e2[u] * hm2[k - u];
    if constexpr (D >= 3) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k - u >= 0 && k - u <= D - 3) s = s + f.e3[u] * hm3[k - u];
    }
    h[k] = s;
  }
}
template <int D, int DMAX, typename F>
__device__ __forceinline__ void moment_degrees_from_dd(const TriFormsDD& tf, const dd* hm3, const dd* hm2,
                                                       const dd* hm1, F& fn) {
  dd h[D + 1], m[D + 1];
  h_next_dd<D>(tf, hm2, hm3, h);
#pragma unroll
  for (int b = 0; b <= D; ++b) m[b] = h[b] * (mkdd(2.0) / mkdd(double((D + 1) * (D + 2)) * binom(D, b)));
  fn(Deg<D>{}, m);
  if constexpr (D < DMAX) moment_degrees_from_dd<D + 1, DMAX>(tf, hm2, hm1, h, fn);
}
template <int DMAX, int P, typename F>
__device__ __forceinline__ void moment_degrees_dd(const GeoDD<P>& g, F&& fn) {
  const TriFormsDD tf = tri_forms_dd<P>(g);
  const dd h0[1] = {mkdd(1.0)}, h1[2] = {mkdd(0.0), mkdd(0.0)};
  fn(Deg<0>{}, h0);
  if constexpr (DMAX >= 1) fn(Deg<1>{}, h1);
  if constexpr (DMAX >= 2) moment_degrees_from_dd<2, DMAX>(tf, nullptr, h0, h1, fn);
}
// lump_correct in double-double
template <int P>
__device__ __forceinline__ void lump_correct_dd(const GeoDD<P>& g, const dd (&x)[Dims<P>::N], dd (&lim)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N;
  dd dj[N];
  dj[0] = mkdd(0.0);
#pragma unroll
  for (int j = 1; j < N; ++j) dj[j] = (x[j] - lim[j]) / mkdd(tfab(j));
  dd mlo[N], s[N], den[N];
#pragma unroll
  for (int i = 0; i < N; ++i) s[i] = den[i] = mkdd(0.0);
  moment_degrees_dd<2 * P>(g, [&](auto dc, const dd* m) {
    constexpr int D = decltype(dc)::value;
    if constexpr (D <= P) {
#pragma unroll
      for (int b = 0; b <= D; ++b) mlo[D * (D + 1) / 2 + b] = m[b];
    }
    if constexpr (D >= 2) {
#pragma unroll
      for (int i = 1; i < N; ++i) {
#pragma unroll
        for (int j = 1; j < N; ++j)
          if (tdeg(i) + tdeg(j) == D) s[i] = s[i] + m[tb(i) + tb(j)] * dj[j];
        if (2 * tdeg(i) == D) den[i] = m[2 * tb(i)] - mlo[i] * mlo[i];
      }
    }
  });
  dd S0 = mkdd(0.0);
#pragma unroll
  for (int j = 1; j < N; ++j) S0 = S0 + mlo[j] * dj[j];
#pragma unroll
  for (int i = 1; i < N; ++i) lim[i] = lim[i] + (s[i] - mlo[i] * S0) * mkdd(tfab(i)) / den[i];
}

// Occupancy floor of the limiter at p = 3, 4 (waves per SIMD): the register
// allocator needs the hint to stay within 168 / 256 registers there.
template <int P>
constexpr int limit_min_waves() { return P == 3 ? 3 : P == 4 ? 2 : 1; }

// Limiting factor of one vertex, core/computeVertexBasedLimiter.m:89-105:
// the deviation d = val - cen of the vertex value is scaled back into
// [lo - cen, hi - cen], with the reference's tolerance 1e-8 on both sides.
__device__ __forceinline__ double vertex_alpha(double val, double cen, double lo, double hi) {
  const double tol = 1.e-8;
  const double d = val - cen;
  const double dMin = lo - cen;
  const double dMax = hi - cen;
  double al = 1.0;
  if (d < dMin + tol) al = fmax(0.0, fmin(1.0, dMin / (d - tol)));
  if (d > dMax - tol) al = fmax(0.0, fmin(1.0, dMax / (d + tol)));
  return al;
}

template <int P, int MODE, int LT>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(limit_min_waves<P>()))) void limit_kernel(
    LimitArgs a) {
  constexpr int N = Dims<P>::N;
  const int K = a.m.K;
  const int k = elem_index(K, a.m.bs);
  const bool active = k < K;
  const int kk = active ? k : 0;
  const double* __restrict__ T = a.tab;

  double x[N];
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = a.taylor[i * K + kk];
  Geo<P> g;
  g.load(a.m.B, kk, K);
  // Every load that does not depend on the limiting is issued here, before
  // any of it is computed: the vertex ids, then (the 'linear' type) the
  // vertex-patch bounds and Dirichlet values they index, and for the dc/dt
  // limiter the RK operands.  Loaded inside the vertex loop and after the
  // limiting, the compiler waited for each vertex's id and bounds in turn
  // (the loop body branches) and for the RK operands at the end: one memory
  // round trip after another.  Same operations, same bits.
  int vtx[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) vtx[n] = a.m.V0T[n * K + kk];
  [[maybe_unused]] double vlo[3], vhi[3], vdr[3];
  if constexpr (LT == LT_LINEAR) {
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      vlo[n] = a.vmin[vtx[n]];
      vhi[n] = a.vmax[vtx[n]];
    }
    if (a.vertexD) {
#pragma unroll
      for (int n = 0; n < 3; ++n) vdr[n] = a.vertexD[vtx[n]];
    }
  }
  constexpr bool RKOP = MODE != LIM_SOLUTION && P <= 2;  // (p >= 3: the registers are not there)
  [[maybe_unused]] double cs[RKOP ? N : 1], c0v[RKOP ? N : 1], accv[RKOP ? N : 1];
  if constexpr (RKOP) {
    const RKCoef& rk = a.rk;
    if (rk.a1 != 0.0) {
#pragma unroll
      for (int i = 0; i < N; ++i) cs[i] = a.c[i * K + kk];
    }
    if (rk.a0 != 0.0 || (a.accOut && rk.d0 != 0.0)) {
#pragma unroll
      for (int i = 0; i < N; ++i) c0v[i] = a.c0[i * K + kk];
    }
    if (rk.g != 0.0 || (a.accOut && rk.d1 != 0.0)) {
#pragma unroll
      for (int i = 0; i < N; ++i) accv[i] = a.acc[i * K + kk];
    }
  }

  // vertex values of the linear part, core/applySlopeLimiterTaylorLinear.m:117
  // (phi^Taylor_2,3 at vertex n = dX * (x_n - x_c))
  const double ex = (g.B11 + g.B12) / 3.0, ey = (g.B21 + g.B22) / 3.0;
  const double px[3] = {-ex, g.B11 - ex, g.B12 - ex};
  const double py[3] = {-ey, g.B21 - ey, g.B22 - ey};
  const double xc = x[0];
  // patch bounds with the reference's global shift: min over the patch of
  // (c - sNeg), then + sNeg (and max with sPos); the Dirichlet vertex value
  // joins NaN-ignoring afterwards (core/computeMinMaxV0TElementPatch.m:75-82,107-110)
  double alpha = 1.0;
  double lim[N];
  bool changed;
  if constexpr (LT == LT_LINEAR) {
    // core/applySlopeLimiterTaylorLinear.m:108-122
    const double sNeg = fabs(a.gmm[1]) + 1.0;
    const double sPos = fabs(a.gmm[0]) + 1.0;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const double val = x[0] + x[1] * (px[n] * g.dX) + x[2] * (py[n] * g.dX);
      double lo = (vlo[n] - sNeg) + sNeg;
      double hi = (vhi[n] + sPos) - sPos;
      if (a.vertexD) {
        const double dv = vdr[n];
        lo = fmin(lo, dv);
        hi = fmax(hi, dv);
      }
      alpha = fmin(alpha, vertex_alpha(val, xc, lo, hi));
    }
    lim[0] = x[0];
    lim[1] = alpha * x[1];
    lim[2] = alpha * x[2];
    const double keep = alpha == 1.0 ? 1.0 : 0.0;
#pragma unroll
    for (int i = 3; i < N; ++i) lim[i] = keep * x[i];
    changed = alpha != 1.0;
  } else {
    // core/applySlopeLimiterTaylorHierarchicalVertex.m:96-131 and
    // core/applySlopeLimiterTaylorStrict.m:93-127: for ord = p .. 1 and
    // i = 1 .. ord, the reconstruction of the derivative coefficient
    // (ord - i, i - 1) (linear for 'hierarch_vert', of degree p - ord + 1 for
    // 'strict') is limited against the vertex-patch bounds of that
    // coefficient; for ord > 1 the element's own vertex value stands in for
    // the Dirichlet data at Dirichlet vertices.
    const int fl = a.m.flags[kk];
    double xi[3], eta[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      xi[n] = px[n] * g.dX;
      eta[n] = py[n] * g.dX;
    }
    // Taylor basis at the vertices, core/computeTaylorBasesV0T.m:58-63
    // (degree >= 2: (xi^a eta^b - mean) / (a! b!); strict only)
    constexpr int NV = LT == LT_STRICT ? N : 3;
    double phiV[3][NV];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      phiV[n][0] = 1.0;
      phiV[n][1] = xi[n];
      phiV[n][2] = eta[n];
    }
    if constexpr (LT == LT_STRICT && P >= 2) {
      double mlo[N];
      moment_degrees<P>(g, [&](auto dc, const double* m) {
        constexpr int D = decltype(dc)::value;
#pragma unroll
        for (int b = 0; b <= D; ++b) mlo[D * (D + 1) / 2 + b] = m[b];
      });
#pragma unroll
      for (int n = 0; n < 3; ++n) {
#pragma unroll
        for (int m = 3; m < NV; ++m) {
          const int ea = tdeg(m) - tb(m), eb = tb(m);
          double pw = 1.0;
#pragma unroll
          for (int t = 0; t < ea; ++t) pw *= xi[n];
#pragma unroll
          for (int t = 0; t < eb; ++t) pw *= eta[n];
          phiV[n][m] = (pw - mlo[m]) / tfab(m);
        }
      }
    }
    if constexpr (LT == LT_HIERARCH) {
      lim[0] = x[0];
    } else {
#pragma unroll
      for (int m = 0; m < N; ++m) lim[m] = x[m];
    }
    double alphaH = 0.0;
#pragma unroll
    for (int ord = P; ord >= 1; --ord) {
      double alphaOrd = 1.0;
#pragma unroll
      for (int i = 1; i <= ord; ++i) {
        const int f = (ord - 1) * ord / 2 + (i - 1);  // coefficient (ord - i, i - 1)
        const double sNeg = fabs(a.gmm[2 * f + 1]) + 1.0;
        const double sPos = fabs(a.gmm[2 * f]) + 1.0;
        const double cen = LT == LT_STRICT ? lim[f] : x[f];
        double alphaTmp = 1.0;
#pragma unroll
        for (int n = 0; n < 3; ++n) {
          double val;
          if constexpr (LT == LT_HIERARCH) {
            val = x[f] + x[ord * (ord + 1) / 2 + (i - 1)] * xi[n] + x[ord * (ord + 1) / 2 + i] * eta[n];
          } else {
            const int pR = P - ord + 1, nI = (pR + 1) * (pR + 2) / 2;
            val = lim[f];
#pragma unroll
            for (int m = 1; m < nI; ++m) {
              const int dm = tdeg(m) + ord - 1, bm = tb(m) + i - 1;
              val += lim[dm * (dm + 1) / 2 + bm] * phiV[n][m];
            }
          }
          const int v = vtx[n];
          double lo = (a.vmin[size_t(f) * a.m.numV + v] - sNeg) + sNeg;
          double hi = (a.vmax[size_t(f) * a.m.numV + v] + sPos) - sPos;
          if (ord > 1) {
            if (flag_vdir(fl, n)) {
              lo = fmin(lo, val);
              hi = fmax(hi, val);
            }
          } else if (a.vertexD) {
            const double dv = a.vertexD[v];
            lo = fmin(lo, dv);
            hi = fmax(hi, dv);
          }
          alphaTmp = fmin(alphaTmp, vertex_alpha(val, cen, lo, hi));
        }
        alphaOrd = fmin(alphaOrd, alphaTmp);
      }
      if constexpr (LT == LT_HIERARCH) {
        alphaH = fmax(alphaH, alphaOrd);
#pragma unroll
        for (int m = ord * (ord + 1) / 2; m < (ord + 1) * (ord + 2) / 2; ++m) lim[m] = alphaH * x[m];
        alpha = alphaH;
      } else {
#pragma unroll
        for (int m = ord * (ord + 1) / 2; m < N; ++m) lim[m] = alphaOrd * lim[m];
        alpha = alphaOrd;
      }
    }
    changed = false;
#pragma unroll
    for (int m = 1; m < N; ++m) changed |= lim[m] != x[m];
  }

  if constexpr (MODE != LIM_SOLUTION && P >= 2) {
    // selective mass lumping: lim + diag(MT)^-1 MT (orig - lim)
    if (changed) lump_correct<P>(g, x, lim);
  }
  double y[N];
  taylor_to_dg<P>(lim, g, const_tab(T + a.off.FinvR), y);

  if (a.alphaOut && active) a.alphaOut[k] = alpha;
  if constexpr (MODE == LIM_SOLUTION) {
    if (active)
#pragma unroll
      for (int i = 0; i < N; ++i) a.out[i * K + k] = y[i];
  } else {
    double nx[N];
    const double tk = a.tauK ? a.tauK[kk] : a.tau;
    if constexpr (RKOP) {
      rk_combine_pre<N>(a.rk, tk, cs, y, c0v, accv, a.accOut, K, kk, active, nx);
    } else {
      double cl[N];
      if (a.rk.a1 != 0.0) {
#pragma unroll
        for (int i = 0; i < N; ++i) cl[i] = a.c[i * K + kk];
      }
      rk_combine<N>(a.rk, tk, cl, y, a.c0, a.acc, a.accOut, K, kk, active, nx);
    }
    if constexpr (MODE == LIM_DOT_RK_DG) {
      if (active)
#pragma unroll
        for (int i = 0; i < N; ++i) a.out[i * K + k] = nx[i];
    } else {
      double Tc[N];
      dg_to_taylor<P>(nx, g, const_tab(T + a.off.FR), Tc);
      if (active) {
#pragma unroll
        for (int i = 0; i < N; ++i) a.out[i * K + k] = Tc[i];
      }
      sink_taylor<P>(Tc, active, a.sink, a.m.bs, elem_blocks(K, a.m.bs));
    }
  }
}

// The 'linear' limiter on DG coefficients (LIM_DOT_RK_DG: dc/dt limiting,
// RK update, new state in DG with its centroid min / max partials;
// LIM_SOLUTION_DG: the state limited in place).  Its decision needs only the
// Taylor coefficients of degree <= 1, which dg_to_taylor_lin forms with the
// operations of dg_to_taylor; a limited element then takes limit_kernel's
// path with the same operations, and an element the limiter keeps
// stays as it was -- where the reference maps it to the Taylor basis and back
// (core/applySlopeLimiterDisc.m:56-60), which returns it up to rounding.  So
// the full changes of basis run only in waves that hold a limited element,
// and the solution limiter stores only limited elements.
// FIX (LIM_DOT_RK_DG on a mesh with listed badly shaped elements): the
// listed elements take their limited dc/dt from limit_dot_fixup_kernel.
template <int P, int MODE, bool FIX = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(limit_min_waves<P>()))) void limit_dg_kernel(
    LimitArgs a) {
  static_assert(MODE == LIM_DOT_RK_DG || MODE == LIM_SOLUTION_DG, "DG-input modes");
  constexpr int N = Dims<P>::N;
  constexpr bool DOT = MODE == LIM_DOT_RK_DG;
  const int K = a.m.K;
  const int k = elem_index(K, a.m.bs);
  const bool active = k < K;
  const int kk = active ? k : 0;
  const double* __restrict__ T = a.tab;
  [[maybe_unused]] const double f00 = T[a.off.FR];  // F_00: the Taylor mean is F_00 x_0

  double xd[N];
#pragma unroll
  for (int i = 0; i < N; ++i) xd[i] = a.taylor[i * K + kk];
  // the geometry only for the dc/dt limiter's full path (mass lumping), loaded there
  [[maybe_unused]] Geo<P> g;
  // every load that does not depend on the limiting up front (as in limit_kernel)
  int vtx[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) vtx[n] = a.m.V0T[n * K + kk];
  double vlo[3], vhi[3], vdr[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    vlo[n] = a.vmin[vtx[n]];
    vhi[n] = a.vmax[vtx[n]];
  }
  if (a.vertexD) {
#pragma unroll
    for (int n = 0; n < 3; ++n) vdr[n] = a.vertexD[vtx[n]];
  }
  // badly shaped element whose limited dc/dt comes from the fix-up kernel
  [[maybe_unused]] const int fi = FIX ? a.fixIdx[kk] : -1;
  constexpr bool RKOP = DOT && P <= 2;
  [[maybe_unused]] double cs[RKOP ? N : 1], c0v[RKOP ? N : 1], accv[RKOP ? N : 1];
  if constexpr (RKOP) {
    const RKCoef& rk = a.rk;
    if (rk.a1 != 0.0) {
#pragma unroll
      for (int i = 0; i < N; ++i) cs[i] = a.c[i * K + kk];
    }
    if (rk.a0 != 0.0 || (a.accOut && rk.d0 != 0.0)) {
#pragma unroll
      for (int i = 0; i < N; ++i) c0v[i] = a.c0[i * K + kk];
    }
    if (rk.g != 0.0 || (a.accOut && rk.d1 != 0.0)) {
#pragma unroll
      for (int i = 0; i < N; ++i) accv[i] = a.acc[i * K + kk];
    }
  }

  // core/applySlopeLimiterTaylorLinear.m:108-122 on the mean and gradient
  double t1[3];
  ref_lin<P>(xd, const_tab(T + a.off.FR), t1);
  const double sNeg = fabs(a.gmm[1]) + 1.0;
  const double sPos = fabs(a.gmm[0]) + 1.0;
  double alpha = 1.0;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const double val = t1[0] + ref_lin_dev(t1, n);
    double lo = (vlo[n] - sNeg) + sNeg;
    double hi = (vhi[n] + sPos) - sPos;
    if (a.vertexD) {
      const double dv = vdr[n];
      lo = fmin(lo, dv);
      hi = fmax(hi, dv);
    }
    alpha = fmin(alpha, vertex_alpha(val, t1[0], lo, hi));
  }
  const bool changed = alpha != 1.0;
  // (every store comes after the last table read: a store before them would
  // turn the tables' scalar loads into vector loads, each waited for alone)

  if constexpr (!DOT) {
    if (changed && active) {
      const double lim[3] = {t1[0], alpha * t1[1], alpha * t1[2]};
      double y[N];
      ref_lin_to_dg<P>(lim, const_tab(T + a.off.FinvR), y);
#pragma unroll
      for (int i = 0; i < N; ++i) a.out[i * K + k] = y[i];
    }
    if (a.alphaOut && active) a.alphaOut[k] = alpha;
  } else {
    double y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = xd[i];
    if (FIX && fi >= 0) {
#pragma unroll
      for (int i = 0; i < N; ++i) y[i] = a.fixY[size_t(i) * a.fixCount + fi];
    } else if (changed) {
      // limit_kernel's path: all Taylor coefficients, the limited ones, the
      // mass lumping, back to DG
      double x[N], lim[N];
      g.load(a.m.B, kk, K);
      dg_to_taylor<P>(xd, g, const_tab(T + a.off.FR), x);
      lim[0] = x[0];
      lim[1] = alpha * x[1];
      lim[2] = alpha * x[2];
#pragma unroll
      for (int i = 3; i < N; ++i) lim[i] = 0.0 * x[i];
      if constexpr (P >= 2) lump_correct<P>(g, x, lim);
      taylor_to_dg<P>(lim, g, const_tab(T + a.off.FinvR), y);
    }
    double nx[N];
    const double tk = a.tauK ? a.tauK[kk] : a.tau;
    if constexpr (RKOP) {
      rk_combine_pre<N>(a.rk, tk, cs, y, c0v, accv, a.accOut, K, kk, active, nx);
    } else {
      double cl[N];
      if (a.rk.a1 != 0.0) {
#pragma unroll
        for (int i = 0; i < N; ++i) cl[i] = a.c[i * K + kk];
      }
      rk_combine<N>(a.rk, tk, cl, y, a.c0, a.acc, a.accOut, K, kk, active, nx);
    }
    if (active) {
#pragma unroll
      for (int i = 0; i < N; ++i) a.out[i * K + k] = nx[i];
    }
    // centroid (Taylor mean) min / max partials: dg_to_taylor's T_0 = F_00 x_0
    double Tc[N];
    Tc[0] = f00 * nx[0];
#pragma unroll
    for (int i = 1; i < N; ++i) Tc[i] = 0.0;
    sink_taylor<P>(Tc, active, a.sink, a.m.bs, elem_blocks(K, a.m.bs));
    if (a.alphaOut && active) a.alphaOut[k] = alpha;
  }
}

// The dc/dt limiter's limited dc/dt (limit_dg_kernel, LIM_DOT_RK_DG) of the
// listed badly shaped elements: alpha with the limiter's own double
// operations, then the full change of basis, the limited coefficients, the
// mass lumping and the way back in double-double (see dd above), rounded
// once at the end.  One thread per listed element.
template <int P>
__global__ __launch_bounds__(BLOCK) void limit_dot_fixup_kernel(LimitArgs a, const int* __restrict__ list, int count,
                                                               double* __restrict__ yOut) {
  constexpr int N = Dims<P>::N;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int K = a.m.K;
  const int k = list[t];
  cdbl* T = const_tab(a.tab);
  double xd[N];
#pragma unroll
  for (int i = 0; i < N; ++i) xd[i] = a.taylor[i * K + k];
  double vlo[3], vhi[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const int v = a.m.V0T[n * K + k];
    vlo[n] = a.vmin[v];
    vhi[n] = a.vmax[v];
  }
  double t1[3];
  ref_lin<P>(xd, T + a.off.FR, t1);
  const double sNeg = fabs(a.gmm[1]) + 1.0;
  const double sPos = fabs(a.gmm[0]) + 1.0;
  double alpha = 1.0;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const double val = t1[0] + ref_lin_dev(t1, n);
    const double lo = (vlo[n] - sNeg) + sNeg;
    const double hi = (vhi[n] + sPos) - sPos;
    alpha = fmin(alpha, vertex_alpha(val, t1[0], lo, hi));
  }
  double y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) y[i] = xd[i];
  if (alpha != 1.0) {
    GeoDD<P> gd;
    gd.load(a.m.B, k, K);
    dd x[N], lim[N], yd[N];
    dg_to_taylor_dd<P>(xd, gd, T + a.off.FR, T + a.off.FloR, x);
    const dd al = mkdd(alpha);
    lim[0] = x[0];
    lim[1] = al * x[1];
    lim[2] = al * x[2];
#pragma unroll
    for (int i = 3; i < N; ++i) lim[i] = mkdd(0.0);
    if constexpr (P >= 2) lump_correct_dd<P>(gd, x, lim);
    taylor_to_dg_dd<P>(lim, gd, T + a.off.FinvR, T + a.off.FinvloR, yd);
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = yd[i].hi + yd[i].lo;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) yOut[size_t(i) * count + t] = y[i];
}

// ------------------------------------------------------ velocity weights --
// The stage kernel's element term needs the velocity only through
//   w_q (B22 u1(q) - B12 u2(q)),  w_q (B11 u2(q) - B21 u1(q))
// at the points of quadRule2D(2p) (core/assembleMatElemDphiPhiFuncDisc.m:196-197
// in quadrature form); formed once per velocity instead of every stage.
template <int P>
__global__ __launch_bounds__(BLOCK) void velocity_weights_kernel(int K, const double* __restrict__ Bp,
                                                                 const double* __restrict__ u1,
                                                                 const double* __restrict__ u2,
                                                                 const double* __restrict__ tab, TabOff off,
                                                                 double* __restrict__ uq) {
  constexpr int N = Dims<P>::N, RE = Dims<P>::RE;
  const int k = gtid();
  if (k >= K) return;
  double a[N], b[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    a[i] = u1[i * K + k];
    b[i] = u2[i * K + k];
  }
  Geo<P> g;
  g.load(Bp, k, K);
  const double* __restrict__ EQ = tab + off.elemQ;
  for (int q = 0; q < RE; ++q) {
    double u1q = 0.0, u2q = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const double ph = EQ[QIdx<P>::phi(q, j)];
      u1q += a[j] * ph;
      u2q += b[j] * ph;
    }
    const double w = EQ[QIdx<P>::w(q)];
    uq[size_t(q) * K + k] = w * (g.B22 * u1q - g.B12 * u2q);
    uq[size_t(RE + q) * K + k] = w * (g.B11 * u2q - g.B21 * u1q);
  }
}

// ------------------------------------------------------------ to_taylor --

template <int P>
__global__ __launch_bounds__(BLOCK) void to_taylor_kernel(ToTaylorArgs a) {
  constexpr int N = Dims<P>::N;
  const int K = a.K;
  const int k = elem_index(K, a.bs);
  const bool active = k < K;
  const int kk = active ? k : 0;
  double x[N], Tc[N];
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = a.c[i * K + kk];
  Geo<P> g;
  g.load(a.B, kk, K);
  dg_to_taylor<P>(x, g, a.tab + a.off.FR, Tc);
  if (active) {
#pragma unroll
    for (int i = 0; i < N; ++i) a.out[i * K + k] = Tc[i];
  }
  sink_taylor<P>(Tc, active, a.sink, a.bs, elem_blocks(K, a.bs));
}

// ---------------------------------------------------- DG-linear limit --
// core/applySlopeLimiterDiscLinear.m:109-123.  Pass 1 (centroid_kernel): the
// centroid values c_1 phi_1(1/3, 1/3) and their min / max partials; pass 2:
// vertex_minmax_kernel on that one field; pass 3 (limit_disc_linear_kernel):
// the vertex values of the DG data's linear part (core/projectDataDisc2DataLagr.m
// at the reference vertices) against the patch bounds, the linear
// coefficients scaled by alpha and the higher ones kept only where alpha = 1.

__global__ __launch_bounds__(BLOCK) void centroid_kernel(int K, const double* __restrict__ c, double phi1c,
                                                         double* __restrict__ cen, MinMaxSink sink) {
  const int k = gtid();
  const bool active = k < K;
  const double v[1] = {active ? c[k] * phi1c : 0.0};
  if (active) cen[k] = v[0];
  sink_taylor<0>(v, active, sink, int(blockDim.x), int(gridDim.x));
}

template <int P>
__global__ __launch_bounds__(BLOCK) void limit_disc_linear_kernel(DiscLinArgs a) {
  constexpr int N = Dims<P>::N;
  const int K = a.m.K;
  const int k = gtid();
  if (k >= K) return;
  double x[N];
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = a.c[i * K + k];
  const double xc = a.cen[k];
  const double sNeg = fabs(a.gmm[1]) + 1.0;
  const double sPos = fabs(a.gmm[0]) + 1.0;
  double alpha = 1.0;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    double val;
    {
      // separately rounded products summed left to right, as
      // core/projectDataDisc2DataLagr.m:53-56 accumulates them
#pragma clang fp contract(off)
      val = x[0] * a.phiV[n][0] + x[1] * a.phiV[n][1] + x[2] * a.phiV[n][2];
    }
    const int v = a.m.V0T[n * K + k];
    double lo = (a.vmin[v] - sNeg) + sNeg;
    double hi = (a.vmax[v] + sPos) - sPos;
    if (a.vertexD) {
      const double dv = a.vertexD[v];
      lo = fmin(lo, dv);
      hi = fmax(hi, dv);
    }
    alpha = fmin(alpha, vertex_alpha(val, xc, lo, hi));
  }
  a.c[K + k] = alpha * x[1];
  a.c[2 * K + k] = alpha * x[2];
  const double keep = alpha == 1.0 ? 1.0 : 0.0;
#pragma unroll
  for (int i = 3; i < N; ++i) a.c[i * K + k] = keep * x[i];
  if (a.alphaOut) a.alphaOut[k] = alpha;
}

// ------------------------------------------------------ setup kernels --
// L2 projection of point values, core/projectFuncCont2DataDisc.m:98-101:
// out = (sum_q f_q w_q phi(q)) / hatM.
template <int P, int R>
__global__ __launch_bounds__(BLOCK) void project_kernel(int K, const double* __restrict__ fq,
                                                        const double* __restrict__ tab, int offW, int offPhi,
                                                        int offMinv, double* __restrict__ out) {
  constexpr int N = Dims<P>::N;
  const int k = gtid();
  if (k >= K) return;
  double r[N];
#pragma unroll
  for (int i = 0; i < N; ++i) r[i] = 0.0;
  for (int q = 0; q < R; ++q) {
    const double f = fq[q * K + k] * tab[offW + q];
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] += f * tab[offPhi + q + i * R];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) s += r[j] * tab[offMinv + j + i * N];
    out[i * K + k] = s;
  }
}

// u . nu at the edge points, core/computeFuncContNuOnQuadEdge.m:61-66
__global__ __launch_bounds__(BLOCK) void vnormal_kernel(int K, int Re, const double* __restrict__ nuE0T,
                                                        const double* __restrict__ u1e,
                                                        const double* __restrict__ u2e, double* __restrict__ vn) {
  const int k = gtid();
  if (k >= K) return;
  for (int n = 0; n < 3; ++n) {
    const double n1 = nuE0T[k + n * K], n2 = nuE0T[k + n * K + 3 * K];
    for (int r = 0; r < Re; ++r) {
      const int idx = (r * 3 + n) * K + k;
      vn[idx] = n1 * u1e[idx] + n2 * u2e[idx];
    }
  }
}

// The stage kernel's edge-flux factors (|E_kn| u.nu) w_r, rounded as the
// product area * v * w_r is (so the flux factor * trace gives the same bits).
__global__ __launch_bounds__(BLOCK) void edge_flux_weights_kernel(int K, int Re, const double* __restrict__ areaE,
                                                                  const double* __restrict__ vn,
                                                                  const double* __restrict__ wEdge,
                                                                  double* __restrict__ vnw) {
  const int k = gtid();
  if (k >= K) return;
  for (int n = 0; n < 3; ++n) {
    const double area = areaE[n * K + k];
    for (int r = 0; r < Re; ++r) {
      const int idx = (r * 3 + n) * K + k;
      vnw[idx] = area * vn[idx] * wEdge[r];
    }
  }
}

// globL - globKD - globKN: detJ hatM fDisc minus the Dirichlet inflow term
// and the Neumann term, advection/preprocessSubStep.m:88-98,
// core/assembleVecEdgePhiIntFuncContVal.m:173-184,
// core/assembleVecEdgePhiIntFuncCont.m:126-135 (with the upwind g_N of
// advection/preprocessSubStep.m:92: g_N where g_N <= 0, else 0)
template <int P>
__global__ __launch_bounds__(BLOCK) void source_kernel(Mesh m, const double* __restrict__ fq,
                                                       const double* __restrict__ cDe, const double* __restrict__ gNe,
                                                       const double* __restrict__ vn,
                                                       const double* __restrict__ tab, TabOff off,
                                                       double* __restrict__ rhs) {
  constexpr int N = Dims<P>::N, RE = Dims<P>::RE, Re = Dims<P>::Re;
  const int K = m.K;
  const int k = gtid();
  if (k >= K) return;
  double r[N];
#pragma unroll
  for (int i = 0; i < N; ++i) r[i] = 0.0;
  if (fq) {
    Geo<P> g;
    g.load(m.B, k, K);
    double pr[N], fd[N];
#pragma unroll
    for (int i = 0; i < N; ++i) pr[i] = 0.0;
    for (int q = 0; q < RE; ++q) {
      const double f = fq[q * K + k] * tab[off.wE + q];
#pragma unroll
      for (int i = 0; i < N; ++i) pr[i] += f * tab[off.phiE + q + i * RE];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += pr[j] * tab[off.hatMinv + j + i * N];
      fd[i] = s;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += tab[off.hatM + i + j * N] * fd[j];
      r[i] = g.det * s;
    }
  }
  if (cDe) {
    const int fl = m.flags[k];
    for (int n = 0; n < 3; ++n) {
      if (!flag_dir(fl, n)) continue;
      const double area = m.areaE[n * K + k];
      for (int r2 = 0; r2 < Re; ++r2) {
        const int idx = (r2 * 3 + n) * K + k;
        const double v = vn[idx];
        const double fx = area * (cDe[idx] * ((v < 0.0) ? v : 0.0)) * tab[off.wEdge + r2];
#pragma unroll
        for (int i = 0; i < N; ++i) r[i] -= fx * tab[off.phi1D + r2 + i * Re + n * Re * N];
      }
    }
  }
  if (gNe) {
    // K_N(k, i) = sum_n [markE0TbdrN] |E_kn| sum_r gN_up(x_knr) w_r phi_i(gamma_n(s_r)),
    // the r-sum first (funcQ0E * (W' .* phi1D), then the edge factor)
    const int fl = m.flags[k];
    for (int n = 0; n < 3; ++n) {
      if (!flag_neu(fl, n)) continue;
      const double area = m.areaE[n * K + k];
      double e[N];
#pragma unroll
      for (int i = 0; i < N; ++i) e[i] = 0.0;
      for (int r2 = 0; r2 < Re; ++r2) {
        const double gn = gNe[(r2 * 3 + n) * K + k];
        const double gu = (gn <= 0.0) ? gn : 0.0 * gn;
#pragma unroll
        for (int i = 0; i < N; ++i) e[i] += gu * (tab[off.wEdge + r2] * tab[off.phi1D + r2 + i * Re + n * Re * N]);
      }
#pragma unroll
      for (int i = 0; i < N; ++i) r[i] -= area * e[i];
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) rhs[i * K + k] = r[i];
}

// per-block partial sums of 2 |T_k| sum_r W_r (c_h - c)^2, core/computeL2Error.m:103-113
template <int P, int R>
__global__ __launch_bounds__(BLOCK) void l2_kernel(int K, const double* __restrict__ B, const double* __restrict__ c,
                                                   const double* __restrict__ cq, const double* __restrict__ tab,
                                                   int offW, int offPhi, double* __restrict__ partial) {
  constexpr int N = Dims<P>::N;
  const int k = gtid();
  double e = 0.0;
  if (k < K) {
    double x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = c[i * K + k];
    for (int q = 0; q < R; ++q) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += x[j] * tab[offPhi + q + j * R];
      const double d = s - cq[q * K + k];
      e += d * d * tab[offW + q];
    }
    const double det = B[k] * B[k + 3 * K] - B[k + 2 * K] * B[k + K];
    e *= det;  // 2 |T_k|
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
  __shared__ double s[BLOCK / 64];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < BLOCK / 64; ++i) t += s[i];
    partial[blockIdx.x] = t;
  }
}


// --------------------------------------------------------------- adjoint --
// Transposed operator per element k (A = G1 + G2 - R):
//   (G^T mu)_j  = sum_q w_q phi_j(q) [a(q) d1 mu(q) + b(q) d2 mu(q)]
//   diagonal R blocks are symmetric (own outflow edges);
//   (R^T mu)_k  += sum over neighbours k' of the block (k', k)^T mu_k',
// gathered from the neighbour's side: its normal velocity, its upwind
// selector and its trace, with the neighbour-trace table indexed by
// (neighbour's local edge, own local edge).  No atomics.
// W = ADJ_STAGE: transposed stage only; ADJ_FUSED: with the stage's velocity
// sensitivities; ADJ_SENS: the sensitivities only; ADJ_SENS_U / ADJ_SENS_V:
// only those w.r.t. u1Disc, u2Disc (element term) / u.nu (edge term).
// UNR: unroll the edge loop (more loads in flight per thread, more registers).
// LATE (the transposed stage alone, p <= 2): the caller has loaded this
// thread's share of the two LDS tables into tv (adj_table_regs) and not yet
// written them; adj_body writes them after issuing the element's own loads
// and waits at a barrier that waits for LDS only, so those loads stay in
// flight across it.  Every thread of the workgroup must then call adj_body;
// inactive ones (active false) pass an element in range and store nothing.
template <int P>
struct AdjTab {
  static constexpr int N1 = Dims<P>::Re * Dims<P>::N * 9, N2 = Dims<P>::Re * Dims<P>::N * 3, TPT = (N1 + N2 + 63) / 64;
};
template <int P>
__device__ __forceinline__ void adj_table_regs(const AdjArgs& a, double (&tv)[AdjTab<P>::TPT]) {
  using AT = AdjTab<P>;
#pragma unroll
  for (int u = 0; u < AT::TPT; ++u) {
    const int i = int(threadIdx.x) + u * a.m.bs;
    tv[u] = i < AT::N1 ? a.tab[a.off.theta + i] : i < AT::N1 + AT::N2 ? a.tab[a.off.phi1D + i - AT::N1] : 0.0;
  }
}

template <int P, int W, bool UNR = false, bool LATE = false>
__device__ __forceinline__ void adj_body(const AdjArgs& a, int k, double* sTheta, double* sPhi1D, bool active = true,
                                         const double* tv = nullptr) {
  constexpr bool ADJ = W == ADJ_STAGE || W == ADJ_FUSED;
  constexpr bool SENS_U = W == ADJ_FUSED || W == ADJ_SENS || W == ADJ_SENS_U;
  constexpr bool SENS_V = W == ADJ_FUSED || W == ADJ_SENS || W == ADJ_SENS_V;
  constexpr bool SENS = SENS_U || SENS_V;
  constexpr int N = Dims<P>::N, RE = Dims<P>::RE, Re = Dims<P>::Re;
  const int K = a.m.K;
  cdbl* T = const_tab(a.tab);  // (scalar loads after the element's stores too)
  double mu[N], u1[N], u2[N], acc[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    mu[i] = a.mu[i * K + k];
    u1[i] = a.u1[i * K + k];
    u2[i] = a.u2[i * K + k];
    acc[i] = 0.0;
  }
  Geo<P> g;
  g.load(a.m.B, k, K);
  const double tk = a.tauK ? a.tauK[k] : a.tau;
  const double al = (1.0 - a.omega) * tk;   // weight of this stage's sensitivity
  // the running c_0 sum of modes 1, 2 with the element's other loads (read
  // where it is used, each load waited alone behind the previous store)
  [[maybe_unused]] double L0v[ADJ ? N : 1];
  if constexpr (ADJ) {
    if (a.mode == 1 || a.mode == 2) {
#pragma unroll
      for (int i = 0; i < N; ++i) L0v[i] = a.L0[i * K + k];
    }
  }
  double c[SENS ? N : 1], gu1[SENS_U ? N : 1], gu2[SENS_U ? N : 1], gvn[SENS_V ? 3 * Re : 1];
  if constexpr (SENS) {
#pragma unroll
    for (int i = 0; i < N; ++i) c[i] = a.c[i * K + k];
  }
  if constexpr (SENS_U) {
#pragma unroll
    for (int i = 0; i < N; ++i) gu1[i] = gu2[i] = 0.0;
  }
  if constexpr (SENS_V) {
#pragma unroll
    for (int i = 0; i < 3 * Re; ++i) gvn[i] = 0.0;
  }
  // The transposed stage alone (ADJ_STAGE) requests the neighbour ids at
  // entry and gathers each neighbour's mu and edge-flux factors one edge
  // ahead (edge 0's while the element term is formed), so that the dependent
  // gathers are in flight behind arithmetic.  A missing neighbour reads the
  // element's own data (in bounds; its contribution is skipped below).
  // (p >= 3 keeps the loop below: the registers are not there)
  constexpr bool FWA = W == ADJ_STAGE && P <= 2;
  constexpr bool kAheadA = FWA;
  [[maybe_unused]] int nbs[3];
  [[maybe_unused]] const int flA = FWA ? a.m.flags[k] : 0;
  double mnb[kAheadA ? 2 : 1][N], vnb[kAheadA ? 2 : 1][Re];
  auto gatherA = [&](int n, double (&dm)[N], double (&dv)[Re]) {
    const int nbk = nbs[n] >= 0 ? nbs[n] : k;
    const int npk = flag_np(flA, n);
#pragma unroll
    for (int i = 0; i < N; ++i) dm[i] = a.mu[i * K + nbk];
#pragma unroll
    for (int r = 0; r < Re; ++r) dv[r] = a.vnw[(r * 3 + npk) * K + nbk];
  };
  if constexpr (FWA) {
#pragma unroll
    for (int n = 0; n < 3; ++n) nbs[n] = a.m.nbr[n * K + k];
    if constexpr (kAheadA) gatherA(0, mnb[0], vnb[0]);
  }
  if constexpr (LATE) {
    using AT = AdjTab<P>;
#pragma unroll
    for (int u = 0; u < AT::TPT; ++u) {
      const int i = int(threadIdx.x) + u * a.m.bs;
      if (i < AT::N1) sTheta[i] = tv[u];
      else if (i < AT::N1 + AT::N2) sPhi1D[i - AT::N1] = tv[u];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
  if constexpr (P >= 1 && (ADJ || SENS_U)) {
    constexpr int RG = Dims<P>::RG, UQ = RG <= 6 ? RG : 1;
#pragma unroll UQ
    for (int q = 0; q < RG; ++q) {
      double m1 = 0.0, m2 = 0.0, u1q = 0.0, u2q = 0.0, cq = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const double ph = T[a.off.elemG + QIdx<P>::phi(q, j)];
        m1 += mu[j] * T[a.off.elemG + QIdx<P>::d1(q, j)];
        m2 += mu[j] * T[a.off.elemG + QIdx<P>::d2(q, j)];
        u1q += u1[j] * ph;
        u2q += u2[j] * ph;
        if constexpr (SENS_U) cq += c[j] * ph;
      }
      const double wq = T[a.off.elemG + QIdx<P>::w(q)];
      if constexpr (ADJ) {
        const double f = wq * ((g.B22 * u1q - g.B12 * u2q) * m1 + (g.B11 * u2q - g.B21 * u1q) * m2);
#pragma unroll
        for (int j = 0; j < N; ++j) acc[j] += T[a.off.elemG + QIdx<P>::phi(q, j)] * f;
      }
      if constexpr (SENS_U) {
        const double wc = al * wq * cq;
        const double f1 = wc * (g.B22 * m1 - g.B21 * m2);
        const double f2 = wc * (g.B11 * m2 - g.B12 * m1);
#pragma unroll
        for (int l = 0; l < N; ++l) {
          gu1[l] += T[a.off.elemG + QIdx<P>::phi(q, l)] * f1;
          gu2[l] += T[a.off.elemG + QIdx<P>::phi(q, l)] * f2;
        }
      }
    }
  }
  if constexpr (FWA) {
    // Same contributions in the same order as the loop below; an inactive
    // point adds T * 0 (acc unchanged) instead of branching, since the
    // upwind side differs from lane to lane.
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const int b = kAheadA ? (n & 1) : 0;
      if constexpr (kAheadA) {
        if (n < 2) gatherA(n + 1, mnb[b ^ 1], vnb[b ^ 1]);
      } else {
        gatherA(n, mnb[0], vnb[0]);
      }
      const int nb = nbs[n];
      const int np = flag_np(flA, n);
      const bool outOn = flag_out(flA, n);
#pragma unroll
      for (int r = 0; r < Re; ++r) {
        const double v = a.vnw[(r * 3 + n) * K + k];
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) t += mu[i] * T[a.off.edgeQ + QIdx<P>::ephi(n, r, i)];
        const double fx = (v > 0.0 && outOn) ? v * t : 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) acc[j] -= T[a.off.edgeQ + QIdx<P>::ephi(n, r, j)] * fx;
      }
      if (nb >= 0) {
#pragma unroll
        for (int r = 0; r < Re; ++r) {
          const double v = vnb[b][r];
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) t += mnb[b][i] * sPhi1D[r + i * Re + np * Re * N];
          const double fx = v < 0.0 ? v * t : 0.0;
          const double* th = sTheta + r + np * Re * N + n * Re * N * 3;
#pragma unroll
          for (int j = 0; j < N; ++j) acc[j] -= th[j * Re] * fx;
        }
      }
    }
  }
  const int fl = a.m.flags[k];
  constexpr int kEdgeUnroll = UNR ? 3 : 1;
#pragma unroll kEdgeUnroll
  for (int n = 0; n < (!FWA && (ADJ || SENS_V) ? 3 : 0); ++n) {
    // the transposed stage alone reads the edge-flux factors |E| w_r vn
    // (the same products as area * v * w_r below)
    constexpr bool FW = W == ADJ_STAGE;
    const double area = FW ? 1.0 : a.m.areaE[n * K + k];
    const int nb = a.m.nbr[n * K + k];
    const int np = flag_np(fl, n);  // neighbour's local edge
    const bool outOn = flag_out(fl, n);
    // own edge points: symmetric outflow block, and the sensitivities
#pragma unroll
    for (int r = 0; r < Re; ++r) {
      const double v = FW ? a.vnw[(r * 3 + n) * K + k] : a.vn[(r * 3 + n) * K + k];
      const bool up = v > 0.0 && outOn;
      const bool in = v < 0.0 && nb >= 0;
      if ((ADJ && up) || (SENS_V && (up || in))) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) t += mu[i] * T[a.off.edgeQ + QIdx<P>::ephi(n, r, i)];
        const double wr = T[a.off.edgeQ + QIdx<P>::ew(n, r)];
        if (ADJ && up) {
          const double fx = FW ? v * t : area * v * wr * t;
#pragma unroll
          for (int j = 0; j < N; ++j) acc[j] -= T[a.off.edgeQ + QIdx<P>::ephi(n, r, j)] * fx;
        }
        if constexpr (SENS_V) {
          double cu = 0.0;
          if (up) {
#pragma unroll
            for (int j = 0; j < N; ++j) cu += c[j] * T[a.off.edgeQ + QIdx<P>::ephi(n, r, j)];
          } else {
            const double* th = sTheta + r + n * Re * N + np * Re * N * 3;
#pragma unroll
            for (int j = 0; j < N; ++j) cu += a.c[j * K + nb] * th[j * Re];
          }
          gvn[r * 3 + n] -= al * area * wr * t * cu;
        }
      }
    }
    // transposed off-diagonal block: the neighbour's inflow through this edge
    if (ADJ && nb >= 0) {
      double mn[N];
#pragma unroll
      for (int i = 0; i < N; ++i) mn[i] = a.mu[i * K + nb];
#pragma unroll
      for (int r = 0; r < Re; ++r) {
        // the neighbour's factor carries the same edge length
        const double v = FW ? a.vnw[(r * 3 + np) * K + nb] : a.vn[(r * 3 + np) * K + nb];
        if (v < 0.0) {
          double t = 0.0;
#pragma unroll
          for (int i = 0; i < N; ++i) t += mn[i] * sPhi1D[r + i * Re + np * Re * N];
          const double fx = FW ? v * t : area * v * T[a.off.edgeQ + QIdx<P>::ew(n, r)] * t;
          const double* th = sTheta + r + np * Re * N + n * Re * N * 3;
#pragma unroll
          for (int j = 0; j < N; ++j) acc[j] -= th[j * Re] * fx;
        }
      }
    }
  }
  if constexpr (SENS_U) {
#pragma unroll
    for (int l = 0; l < N; ++l) {
      a.Gu1[l * K + k] += gu1[l];
      a.Gu2[l * K + k] += gu2[l];
    }
  }
  if constexpr (SENS_V) {
#pragma unroll
    for (int i = 0; i < 3 * Re; ++i) a.Gvn[i * K + k] += gvn[i];
  }
  if constexpr (!ADJ) return;
  double z[N];
  mass_solve<P>(T + a.off.hatMinvR, acc, 1.0 / g.det, z);
  if (!active) return;
  if (a.mode == 4) {
#pragma unroll
    for (int i = 0; i < N; ++i) a.out[i * K + k] = z[i];
    return;
  }
  const double w1 = 1.0 - a.omega;
  if (a.mode == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      a.out[i * K + k] = w1 * (mu[i] + tk * z[i]);
      a.L0[i * K + k] = a.omega * mu[i];
    }
  } else if (a.mode == 1) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      a.out[i * K + k] = w1 * (mu[i] + tk * z[i]);
      a.L0[i * K + k] = L0v[i] + a.omega * mu[i];
    }
  } else if (a.mode == 2) {
#pragma unroll
    for (int i = 0; i < N; ++i) a.out[i * K + k] = L0v[i] + a.omega * mu[i] + w1 * (mu[i] + tk * z[i]);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) a.out[i * K + k] = a.omega * mu[i] + w1 * (mu[i] + tk * z[i]);
  }
}

template <int P>
__device__ __forceinline__ void adj_tables(const AdjArgs& a, double* sTheta, double* sPhi1D) {
  constexpr int N = Dims<P>::N, Re = Dims<P>::Re;
  fill_lds(sTheta, a.tab + a.off.theta, Re * N * 9, int(blockDim.x));
  fill_lds(sPhi1D, a.tab + a.off.phi1D, Re * N * 3, int(blockDim.x));
  __syncthreads();
}

// One thread per element.
template <int P, int W>
__global__ __launch_bounds__(BLOCK) void adj_stage_kernel(AdjArgs a) {
  constexpr int N = Dims<P>::N, Re = Dims<P>::Re;
  __shared__ double sTheta[Re * N * 9];
  __shared__ double sPhi1D[Re * N * 3];
  if constexpr (W == ADJ_STAGE && P <= 2) {
    // the tables are written to LDS inside adj_body, after the element's loads
    double tv[AdjTab<P>::TPT];
    adj_table_regs<P>(a, tv);
    const int k = elem_index(a.m.K, a.m.bs);
    const bool active = k < a.m.K;
    adj_body<P, W, false, true>(a, active ? k : 0, sTheta, sPhi1D, active, tv);
  } else {
    adj_tables<P>(a, sTheta, sPhi1D);
    const int k = gtid();
    if (k >= a.m.K) return;
    adj_body<P, W>(a, k, sTheta, sPhi1D);
  }
}

// Three threads per element: the first third of the workgroup applies the
// transposed stage to E elements, the second accumulates the same elements'
// sensitivities w.r.t. u1Disc, u2Disc, the third those w.r.t. u.nu.  They
// share only read-only inputs and write disjoint outputs, so they need no
// synchronisation; on small meshes this triples the waves in flight.  Each
// output is formed by the same operations as in the one-thread variant.
template <int P>
__global__ __launch_bounds__(BLOCK) void adj_split_kernel(AdjArgs a) {
  constexpr int N = Dims<P>::N, Re = Dims<P>::Re;
  __shared__ double sTheta[Re * N * 9];
  __shared__ double sPhi1D[Re * N * 3];
  adj_tables<P>(a, sTheta, sPhi1D);
  const int E = int(blockDim.x) / 3;
  const int t = threadIdx.x;
  const int role = t / E;
  const int k = logical_block() * E + (t - role * E);
  if (k >= a.m.K) return;
  if (role == 0) adj_body<P, ADJ_STAGE, true>(a, k, sTheta, sPhi1D);
  else if (role == 1) adj_body<P, ADJ_SENS_U, true>(a, k, sTheta, sPhi1D);
  else adj_body<P, ADJ_SENS_V, true>(a, k, sTheta, sPhi1D);
}

// ----------------------------------------- deferred velocity sensitivities --
// One adjoint stage's sensitivity contributions, formed from zero with the
// operations of adj_body's SENS_U / SENS_V parts (so that adding them to the
// running sums in the same order gives the same bits).
template <int P, class TabPtr>
__device__ __forceinline__ void sens_elem_stage(TabPtr T, const TabOff& off, const Geo<P>& g,
                                                const double (&mu)[Dims<P>::N], const double (&c)[Dims<P>::N],
                                                double al, double (&gu1)[Dims<P>::N], double (&gu2)[Dims<P>::N]) {
  constexpr int N = Dims<P>::N, RG = Dims<P>::RG;
#pragma unroll
  for (int l = 0; l < N; ++l) gu1[l] = gu2[l] = 0.0;
  constexpr int UQ = RG <= 6 ? RG : 1;
#pragma unroll UQ
  for (int q = 0; q < RG; ++q) {
    double m1 = 0.0, m2 = 0.0, cq = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const double ph = T[off.elemG + QIdx<P>::phi(q, j)];
      m1 += mu[j] * T[off.elemG + QIdx<P>::d1(q, j)];
      m2 += mu[j] * T[off.elemG + QIdx<P>::d2(q, j)];
      cq += c[j] * ph;
    }
    const double wq = T[off.elemG + QIdx<P>::w(q)];
    const double wc = al * wq * cq;
    const double f1 = wc * (g.B22 * m1 - g.B21 * m2);
    const double f2 = wc * (g.B11 * m2 - g.B12 * m1);
#pragma unroll
    for (int l = 0; l < N; ++l) {
      gu1[l] += T[off.elemG + QIdx<P>::phi(q, l)] * f1;
      gu2[l] += T[off.elemG + QIdx<P>::phi(q, l)] * f2;
    }
  }
}

// Velocity sensitivities of one pass over consecutive adjoint stages (sweep
// order), in chunks of one time step: a chunk's stage contributions are
// summed from zero in sweep order and the chunk sum is added to the running
// G (loaded at entry, written once at the end), so the bits do not depend on
// how the stages are split into passes (the checkpoint interval).  Role 0
// (the first wave of a workgroup) forms the element-term ones (G_u1, G_u2),
// role 1 the edge-term ones (G_vn), for the same 64 elements; role 0 loads
// the next stage's inputs while it forms the current one.  Each thread runs
// through the whole pass: on small meshes the pass holds one wave per SIMD
// and runs beside the reverse sweep's kernels instead of crowding them out
// (a grid of one thread per element and chunk, with the partial sums folded
// afterwards, took the whole GPU for 0.24 ms per 20 steps at K = 32 768 and
// stalled the sweep meanwhile: gradient 3.04 -> 2.90x the forward pass).
template <int P>
__global__ __launch_bounds__(BLOCK) void sens_seq_kernel(SensArgs a) {
  constexpr int N = Dims<P>::N, Re = Dims<P>::Re;
  __shared__ double sTheta[Re * N * 9];
  fill_lds(sTheta, a.tab + a.off.theta, Re * N * 9, int(blockDim.x));
  __syncthreads();
  constexpr int E = 64;
  const int role = threadIdx.x / E;
  const int k = logical_block() * E + (threadIdx.x - role * E);
  const int K = a.m.K;
  if (k >= K) return;
  const int len = a.count / a.nchunk;  // stages per chunk (one time step)
  const int e0 = a.first, e1 = a.first + a.count;
  // the tables through the constant address space: with the opaque offsets
  // below (an asm statement, which the compiler takes as a possible memory
  // write), plain loads became vector loads waited for in small batches
  cdbl* T = const_tab(a.tab);
  Geo<P> g;
  g.load(a.m.B, k, K);
  const double tk = a.tauK ? a.tauK[k] : a.tau;
  if (role == 0) {
    double G1[N], G2[N], P1[N], P2[N], mu[N], c[N];
#pragma unroll
    for (int l = 0; l < N; ++l) {
      G1[l] = a.Gu1[l * K + k];
      G2[l] = a.Gu2[l * K + k];
      P1[l] = P2[l] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      mu[i] = a.mus[e0][i * K + k];
      c[i] = a.cs[e0][i * K + k];
    }
#pragma unroll 1
    for (int e = e0; e < e1; ++e) {
      double mun[N], cnx[N];
      const bool more = e + 1 < e1;
      if (more) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          mun[i] = a.mus[e + 1][i * K + k];
          cnx[i] = a.cs[e + 1][i * K + k];
        }
      }
      const double al = (1.0 - a.omegas[e]) * tk;
      // the table reads stay inside the loop (hoisted, they would not fit
      // in registers): an opaque zero offset per iteration
      int tz = 0;
      asm volatile("" : "+s"(tz));
      double g1[N], g2[N];
      sens_elem_stage<P>(T + tz, a.off, g, mu, c, al, g1, g2);
#pragma unroll
      for (int l = 0; l < N; ++l) {
        P1[l] += g1[l];
        P2[l] += g2[l];
      }
      if ((e - e0 + 1) % len == 0) {  // end of a chunk: add its sum
#pragma unroll
        for (int l = 0; l < N; ++l) {
          G1[l] += P1[l];
          G2[l] += P2[l];
          P1[l] = P2[l] = 0.0;
        }
      }
      if (more) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          mu[i] = mun[i];
          c[i] = cnx[i];
        }
      }
    }
#pragma unroll
    for (int l = 0; l < N; ++l) {
      a.Gu1[l * K + k] = G1[l];
      a.Gu2[l * K + k] = G2[l];
    }
  } else {
    const int fl = a.m.flags[k];
    int nb[3];
    double area[3], v[3][Re];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      nb[n] = a.m.nbr[n * K + k];
      area[n] = a.m.areaE[n * K + k];
#pragma unroll
      for (int r = 0; r < Re; ++r) v[n][r] = a.vn[(r * 3 + n) * K + k];
    }
    double G[3 * Re], Pc[3 * Re];
#pragma unroll
    for (int i = 0; i < 3 * Re; ++i) {
      G[i] = a.Gvn[i * K + k];
      Pc[i] = 0.0;
    }
#pragma unroll 1
    for (int e = e0; e < e1; ++e) {
      const double* __restrict__ mp = a.mus[e];
      const double* __restrict__ cp = a.cs[e];
      const double al = (1.0 - a.omegas[e]) * tk;
      int tz = 0;
      asm volatile("" : "+s"(tz));
      cdbl* Te = T + tz;
      const double* thb = sTheta + tz;
      double mu[N], c[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        mu[i] = mp[i * K + k];
        c[i] = cp[i * K + k];
      }
      double gv[3 * Re];
#pragma unroll
      for (int i = 0; i < 3 * Re; ++i) gv[i] = 0.0;
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        const int np = flag_np(fl, n);
        const bool outOn = flag_out(fl, n);
        double cn[N];
        const int q = nb[n] >= 0 ? nb[n] : k;
#pragma unroll
        for (int j = 0; j < N; ++j) cn[j] = cp[j * K + q];
#pragma unroll
        for (int r = 0; r < Re; ++r) {
          const bool up = v[n][r] > 0.0 && outOn;
          const bool in = v[n][r] < 0.0 && nb[n] >= 0;
          if (up || in) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) t += mu[i] * Te[a.off.edgeQ + QIdx<P>::ephi(n, r, i)];
            const double wr = Te[a.off.edgeQ + QIdx<P>::ew(n, r)];
            double cu = 0.0;
            if (up) {
#pragma unroll
              for (int j = 0; j < N; ++j) cu += c[j] * Te[a.off.edgeQ + QIdx<P>::ephi(n, r, j)];
            } else {
              const double* th = thb + r + n * Re * N + np * Re * N * 3;
#pragma unroll
              for (int j = 0; j < N; ++j) cu += cn[j] * th[j * Re];
            }
            gv[r * 3 + n] -= al * area[n] * wr * t * cu;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 3 * Re; ++i) Pc[i] += gv[i];
      if ((e - e0 + 1) % len == 0) {
#pragma unroll
        for (int i = 0; i < 3 * Re; ++i) {
          G[i] += Pc[i];
          Pc[i] = 0.0;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 3 * Re; ++i) a.Gvn[i * K + k] = G[i];
  }
}

template <int P>
__global__ __launch_bounds__(BLOCK) void loss_kernel(int K, const double* __restrict__ B, const double* __restrict__ c,
                                                     const double* __restrict__ cObs, const double* __restrict__ tab,
                                                     TabOff off, double* __restrict__ part, double* __restrict__ muT) {
  constexpr int N = Dims<P>::N;
  const int k = gtid();
  double e = 0.0;
  if (k < K) {
    double d[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      d[i] = c[i * K + k] - cObs[i * K + k];
      muT[i * K + k] = 2.0 * d[i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += tab[off.hatM + i + j * N] * d[j];
      e += d[i] * s;
    }
    e *= B[k] * B[k + 3 * K] - B[k + 2 * K] * B[k + K];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  __shared__ double sm[BLOCK / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < BLOCK / 64; ++i) t += sm[i];
    part[blockIdx.x] = t;
  }
}

template <int P>
__global__ __launch_bounds__(BLOCK) void mass_kernel(int K, const double* __restrict__ B, const double* __restrict__ x,
                                                     const double* __restrict__ tab, TabOff off,
                                                     double* __restrict__ out) {
  constexpr int N = Dims<P>::N;
  const int k = gtid();
  if (k >= K) return;
  const double det = B[k] * B[k + 3 * K] - B[k + 2 * K] * B[k + K];
  double v[N];
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = x[i * K + k];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) s += tab[off.hatM + i + j * N] * v[j];
    out[i * K + k] = det * s;
  }
}

// ------------------------------------------------ operator-level applies --
// y = A x for one of the operators the reference assembles as sparse
// matrices, each applied on its own (the stage kernel fuses them):
//   OP_G1, OP_G2, OP_G  the element operators globG{1}, globG{2} and their sum
//                       (core/assembleMatElemDphiPhiFuncDisc.m:181-202) at the
//                       velocity u1Disc, u2Disc, in quadrature form on the 2p rule:
//                       (G1 x)_i = sum_q w_q x(q) u1(q) (B22 d1phi_i - B21 d2phi_i)(q),
//                       (G2 x)_i = sum_q w_q x(q) u2(q) (B11 d2phi_i - B12 d1phi_i)(q)
//   OP_R                the upwind edge operator globR
//                       (core/assembleMatEdgePhiPhiValUpwind.m:261-299): the own
//                       trace where u.nu > 0 on non-Neumann edges, the neighbour's
//                       where u.nu < 0 across interior non-Neumann edges
//   OP_M, OP_MINV       globM = kron(diag(detJ), hatM) (core/assembleMatElemPhiPhi.m:98)
//                       and its block inverse (globM \ x, advection/solveSubStep.m:66)
template <int P>
__global__ __launch_bounds__(BLOCK) void apply_kernel(Mesh m, int op, const double* __restrict__ x,
                                                      const double* __restrict__ u1, const double* __restrict__ u2,
                                                      const double* __restrict__ vnw, const double* __restrict__ tab,
                                                      TabOff off, double* __restrict__ y) {
  constexpr int N = Dims<P>::N, RG = Dims<P>::RG, Re = Dims<P>::Re;
  const int K = m.K;
  const int k = gtid();
  if (k >= K) return;
  double c[N], acc[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    c[i] = x[i * K + k];
    acc[i] = 0.0;
  }
  Geo<P> g;
  g.load(m.B, k, K);
  const double* __restrict__ EQ = tab + off.elemG;  // the element term's rule
  if (op == OP_G || op == OP_G1 || op == OP_G2) {
    const bool w1 = op != OP_G2, w2 = op != OP_G1;
    for (int q = 0; q < RG; ++q) {
      double cq = 0.0, u1q = 0.0, u2q = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const double ph = EQ[QIdx<P>::phi(q, j)];
        cq += c[j] * ph;
        if (w1) u1q += u1[j * K + k] * ph;
        if (w2) u2q += u2[j * K + k] * ph;
      }
      const double wc = EQ[QIdx<P>::w(q)] * cq;
      double f1, f2;
      if (op == OP_G) {
        f1 = wc * (g.B22 * u1q - g.B12 * u2q);
        f2 = wc * (g.B11 * u2q - g.B21 * u1q);
      } else if (op == OP_G1) {
        f1 = wc * (g.B22 * u1q);
        f2 = wc * (-g.B21 * u1q);
      } else {
        f1 = wc * (-g.B12 * u2q);
        f2 = wc * (g.B11 * u2q);
      }
#pragma unroll
      for (int i = 1; i < N; ++i) acc[i] += EQ[QIdx<P>::d1(q, i)] * f1 + EQ[QIdx<P>::d2(q, i)] * f2;
    }
  } else if (op == OP_R) {
    const int fl = m.flags[k];
    for (int n = 0; n < 3; ++n) {
      const int nb = m.nbr[n * K + k];
      const int np = flag_np(fl, n);
      const bool outOn = flag_out(fl, n);
      for (int r = 0; r < Re; ++r) {
        const double v = vnw[(r * 3 + n) * K + k];
        double tr = 0.0;
        if (v > 0.0 && outOn) {
#pragma unroll
          for (int j = 0; j < N; ++j) tr += c[j] * tab[off.edgeQ + QIdx<P>::ephi(n, r, j)];
        } else if (v < 0.0 && nb >= 0) {
          const double* th = tab + off.theta + r + n * Re * N + np * Re * N * 3;
#pragma unroll
          for (int j = 0; j < N; ++j) tr += x[j * K + nb] * th[j * Re];
        }
        const double fx = v * tr;
#pragma unroll
        for (int i = 0; i < N; ++i) acc[i] += tab[off.edgeQ + QIdx<P>::ephi(n, r, i)] * fx;
      }
    }
  } else if (op == OP_M) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += tab[off.hatM + i + j * N] * c[j];
      acc[i] = g.det * s;
    }
  } else {  // OP_MINV
    const double idet = 1.0 / g.det;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < N; ++j) s += tab[off.hatMinvR + i * N + j] * c[j];
      acc[i] = s * idet;
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) y[i * K + k] = acc[i];
}

// core/computeVertexBasedLimiter.m:78-106 on caller data: per element the
// centroid value cen (K), the vertex values valV0T (3 x K) and the Dirichlet
// values dataV0T (3 x K, NaN where markV0TbdrD is false; null: none); the
// raw patch min / max of cen per vertex and their global min / max come from
// vertex_minmax_kernel.  Out: alpha (K), minV0T, maxV0T (3 x K) with the
// reference's global shift (core/computeMinMaxV0TElementPatch.m:75-78,107-110).
__global__ __launch_bounds__(BLOCK) void vertex_limiter_kernel(Mesh m, const double* __restrict__ cen,
                                                               const double* __restrict__ valV0T,
                                                               const double* __restrict__ dataV0T,
                                                               const double* __restrict__ vmin,
                                                               const double* __restrict__ vmax,
                                                               const double* __restrict__ gmm, double* __restrict__ alpha,
                                                               double* __restrict__ minV0T, double* __restrict__ maxV0T) {
  const int K = m.K;
  const int k = gtid();
  if (k >= K) return;
  const double sNeg = fabs(gmm[1]) + 1.0;
  const double sPos = fabs(gmm[0]) + 1.0;
  const double xc = cen[k];
  double al = 1.0;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const int v = m.V0T[n * K + k];
    double lo = (vmin[v] - sNeg) + sNeg;
    double hi = (vmax[v] + sPos) - sPos;
    if (dataV0T) {
      const double dv = dataV0T[n * K + k];
      lo = fmin(lo, dv);
      hi = fmax(hi, dv);
    }
    minV0T[n * K + k] = lo;
    maxV0T[n * K + k] = hi;
    al = fmin(al, vertex_alpha(valV0T[n * K + k], xc, lo, hi));
  }
  alpha[k] = al;
}

__global__ __launch_bounds__(BLOCK) void fill_kernel(size_t n, double v, double* __restrict__ x) {
  const size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (i < n) x[i] = v;
}

// ---------------------------------------------------------- launchers --
// Profiling events of the calling thread: when set, launches record them at
// the kernel's own start and end (hipExtLaunchKernel), so the elapsed time
// is the kernel's duration rather than that of the launch.
static thread_local hipEvent_t t_evStart = nullptr, t_evStop = nullptr;
void set_kernel_events(hipEvent_t start, hipEvent_t stop) {
  t_evStart = start;
  t_evStop = stop;
}
#define FESTUNG_LAUNCH(K_, G_, B_, S_, STREAM_, ...)                                        \
  do {                                                                                      \
    if (t_evStart)                                                                          \
      hipExtLaunchKernelGGL(K_, G_, B_, S_, STREAM_, t_evStart, t_evStop, 0, __VA_ARGS__); \
    else                                                                                    \
      hipLaunchKernelGGL(K_, G_, B_, S_, STREAM_, __VA_ARGS__);                             \
  } while (0)

#define FESTUNG_DISPATCH_P(P_, BODY)                           \
  switch (P_) {                                                \
    case 0: { constexpr int PP = 0; BODY; } break;             \
    case 1: { constexpr int PP = 1; BODY; } break;             \
    case 2: { constexpr int PP = 2; BODY; } break;             \
    case 3: { constexpr int PP = 3; BODY; } break;             \
    case 4: { constexpr int PP = 4; BODY; } break;             \
    default: return hipErrorInvalidValue;                      \
  }

#define FESTUNG_DISPATCH_P1(P_, BODY)                          \
  switch (P_) {                                                \
    case 1: { constexpr int PP = 1; BODY; } break;             \
    case 2: { constexpr int PP = 2; BODY; } break;             \
    case 3: { constexpr int PP = 3; BODY; } break;             \
    case 4: { constexpr int PP = 4; BODY; } break;             \
    default: return hipErrorInvalidValue;                      \
  }

static inline dim3 grid_for(int n) { return dim3((n + BLOCK - 1) / BLOCK); }
static inline dim3 grid_for(int n, int bs) { return dim3((n + bs - 1) / bs); }
bool valid_elem_block(int bs) { return bs == 64 || bs == 128 || bs == 256; }

hipError_t launch_stage(int p, int mode, const StageArgs& a, hipStream_t s) {
  const int BS = a.m.bs;
  if (!valid_elem_block(BS)) return hipErrorInvalidValue;
  dim3 gr = grid_for(a.m.K, BS);
  // dynamic LDS: the MFMA element term's hand-over buffer, one per wave
  const size_t lds = mfma_elem(p) ? size_t(BS / 64) * MfmaElem<4>::xchg * sizeof(double) : 0;
#define L_STAGE(MODE_)                                                                    \
  FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((stage_kernel<PP, MODE_>), gr, dim3(BS), lds, s, a))
#define L_STAGE1(MODE_)                                                                   \
  FESTUNG_DISPATCH_P1(p, FESTUNG_LAUNCH((stage_kernel<PP, MODE_>), gr, dim3(BS), lds, s, a))
  switch (mode) {
    case OUT_RK_DG: L_STAGE(OUT_RK_DG); break;
    case OUT_RK_TAYLOR: L_STAGE1(OUT_RK_TAYLOR); break;
    case OUT_DOT_TAYLOR: L_STAGE1(OUT_DOT_TAYLOR); break;
    case OUT_DOT_DG: L_STAGE(OUT_DOT_DG); break;
    case OUT_DOT_DG_CEN: L_STAGE1(OUT_DOT_DG_CEN); break;
    default: return hipErrorInvalidValue;
  }
#undef L_STAGE
#undef L_STAGE1
  return hipGetLastError();
}

hipError_t launch_edge_flux_weights(int K, int Re, const double* areaE, const double* vn, const double* wEdge,
                                    double* vnw, hipStream_t s) {
  FESTUNG_LAUNCH(edge_flux_weights_kernel, grid_for(K), dim3(BLOCK), 0, s, K, Re, areaE, vn, wEdge, vnw);
  return hipGetLastError();
}

hipError_t launch_vertex(const VertexArgs& a, hipStream_t s) {
  // workgroup 0 folds the partials, the others take the vertices
  FESTUNG_LAUNCH(vertex_minmax_kernel, dim3(grid_for(a.numV).x + 1), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_limit(int p, int mode, int type, const LimitArgs& a, hipStream_t s) {
  const int BS = a.m.bs;
  if (p < 1 || !valid_elem_block(BS)) return hipErrorInvalidValue;
  dim3 gr = grid_for(a.m.K, BS);
#define L_LIM(MODE_, LT_) \
  FESTUNG_DISPATCH_P1(p, FESTUNG_LAUNCH((limit_kernel<PP, MODE_, LT_>), gr, dim3(BS), 0, s, a))
#define L_LIM_T(MODE_)                                 \
  switch (type) {                                      \
    case LT_LINEAR: L_LIM(MODE_, LT_LINEAR); break;     \
    case LT_HIERARCH: L_LIM(MODE_, LT_HIERARCH); break; \
    case LT_STRICT: L_LIM(MODE_, LT_STRICT); break;     \
    default: return hipErrorInvalidValue;              \
  }
#define L_LIM_DG(MODE_, FIX_) \
  FESTUNG_DISPATCH_P1(p, FESTUNG_LAUNCH((limit_dg_kernel<PP, MODE_, FIX_>), gr, dim3(BS), 0, s, a))
  switch (mode) {
    case LIM_SOLUTION: L_LIM_T(LIM_SOLUTION); break;
    case LIM_DOT_RK_TAYLOR: L_LIM_T(LIM_DOT_RK_TAYLOR); break;
    case LIM_DOT_RK_DG:
      if (type != LT_LINEAR || a.sink.F != 1 || (a.fixIdx && (!a.fixY || a.fixCount <= 0)))
        return hipErrorInvalidValue;
      if (a.fixIdx) {
        L_LIM_DG(LIM_DOT_RK_DG, true);
      } else {
        L_LIM_DG(LIM_DOT_RK_DG, false);
      }
      break;
    case LIM_SOLUTION_DG:
      if (type != LT_LINEAR || a.out != a.taylor) return hipErrorInvalidValue;
      L_LIM_DG(LIM_SOLUTION_DG, false);
      break;
    default: return hipErrorInvalidValue;
  }
#undef L_LIM_DG
#undef L_LIM
#undef L_LIM_T
  return hipGetLastError();
}

hipError_t launch_limit_dot_fixup(int p, const LimitArgs& a, const int* list, int count, double* y, hipStream_t s) {
  if (p < 2 || count <= 0) return hipErrorInvalidValue;
  FESTUNG_DISPATCH_P1(p, FESTUNG_LAUNCH((limit_dot_fixup_kernel<PP>), grid_for(count), dim3(BLOCK), 0, s, a, list,
                                        count, y));
  return hipGetLastError();
}

hipError_t launch_to_taylor(int p, const ToTaylorArgs& a, hipStream_t s) {
  const int BS = a.bs;
  if (!valid_elem_block(BS)) return hipErrorInvalidValue;
  dim3 gr = grid_for(a.K, BS);
  FESTUNG_DISPATCH_P1(p, FESTUNG_LAUNCH((to_taylor_kernel<PP>), gr, dim3(BS), 0, s, a));
  return hipGetLastError();
}

hipError_t launch_centroid(int K, int bs, const double* c, double phi1c, double* cen, const MinMaxSink& sink,
                           hipStream_t s) {
  if (!valid_elem_block(bs)) return hipErrorInvalidValue;
  FESTUNG_LAUNCH(centroid_kernel, grid_for(K, bs), dim3(bs), 0, s, K, c, phi1c, cen, sink);
  return hipGetLastError();
}

hipError_t launch_limit_disc_linear(int p, const DiscLinArgs& a, hipStream_t s) {
  const int BS = a.m.bs;
  if (!valid_elem_block(BS)) return hipErrorInvalidValue;
  dim3 gr = grid_for(a.m.K, BS);
  FESTUNG_DISPATCH_P1(p, FESTUNG_LAUNCH((limit_disc_linear_kernel<PP>), gr, dim3(BS), 0, s, a));
  return hipGetLastError();
}

hipError_t launch_velocity_weights(int p, int K, const double* B, const double* u1, const double* u2, const double* tab,
                                   const TabOff& off, double* uq, hipStream_t s) {
  switch (p) {
    case 1: FESTUNG_LAUNCH(velocity_weights_kernel<1>, grid_for(K), dim3(BLOCK), 0, s, K, B, u1, u2, tab, off, uq); break;
    case 2: FESTUNG_LAUNCH(velocity_weights_kernel<2>, grid_for(K), dim3(BLOCK), 0, s, K, B, u1, u2, tab, off, uq); break;
    case 3: FESTUNG_LAUNCH(velocity_weights_kernel<3>, grid_for(K), dim3(BLOCK), 0, s, K, B, u1, u2, tab, off, uq); break;
    default: return hipSuccess;
  }
  return hipGetLastError();
}

hipError_t launch_project(int p, bool ordF, int K, const double* fq, const double* tab, const TabOff& off,
                          double* out, hipStream_t s) {
  dim3 gr = grid_for(K);
  if (ordF) {
    FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((project_kernel<PP, Dims<PP>::RF>), gr, dim3(BLOCK), 0, s, K, fq, tab,
                                             off.wF, off.phiF, off.hatMinv, out));
  } else {
    FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((project_kernel<PP, Dims<PP>::RE>), gr, dim3(BLOCK), 0, s, K, fq, tab,
                                             off.wE, off.phiE, off.hatMinv, out));
  }
  return hipGetLastError();
}

hipError_t launch_vnormal(int K, int Re, const double* nuE0T, const double* u1e, const double* u2e, double* vn,
                          hipStream_t s) {
  FESTUNG_LAUNCH(vnormal_kernel, grid_for(K), dim3(BLOCK), 0, s, K, Re, nuE0T, u1e, u2e, vn);
  return hipGetLastError();
}

hipError_t launch_source(int p, const Mesh& m, const double* fq, const double* cDe, const double* gNe,
                         const double* vn, const double* tab, const TabOff& off, double* rhs, hipStream_t s) {
  const int BS = m.bs;
  if (!valid_elem_block(BS)) return hipErrorInvalidValue;
  dim3 gr = grid_for(m.K, BS);
  FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((source_kernel<PP>), gr, dim3(BS), 0, s, m, fq, cDe, gNe, vn, tab, off, rhs));
  return hipGetLastError();
}

hipError_t launch_l2(int p, bool ordF, int K, const double* B, const double* c, const double* cq, const double* tab,
                     const TabOff& off, double* partial, hipStream_t s) {
  dim3 gr = grid_for(K);
  if (ordF) {
    FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((l2_kernel<PP, Dims<PP>::RF>), gr, dim3(BLOCK), 0, s, K, B, c, cq, tab,
                                             off.wF, off.phiF, partial));
  } else {
    FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((l2_kernel<PP, Dims<PP>::RE>), gr, dim3(BLOCK), 0, s, K, B, c, cq, tab,
                                             off.wE, off.phiE, partial));
  }
  return hipGetLastError();
}

// dst[d + j K] = src[idx[d] + j K], j < cols (reorder a K x cols array)
__global__ __launch_bounds__(BLOCK) void permute_kernel(int K, int cols, const double* __restrict__ src,
                                                        const int* __restrict__ idx, double* __restrict__ dst) {
  const int d = gtid();
  if (d >= K) return;
  const int s = idx[d];
  for (int j = 0; j < cols; ++j) dst[d + size_t(j) * K] = src[s + size_t(j) * K];
}

hipError_t launch_permute(int K, int cols, const double* src, const int* idx, double* dst, hipStream_t s) {
  FESTUNG_LAUNCH(permute_kernel, grid_for(K), dim3(BLOCK), 0, s, K, cols, src, idx, dst);
  return hipGetLastError();
}

__global__ __launch_bounds__(BLOCK) void rdet_kernel(int K, const double* __restrict__ B, double* __restrict__ rdet) {
  const int k = gtid();
  if (k >= K) return;
  Geo<1> g;
  g.load(B, k, K);
  rdet[k] = 1.0 / g.det;
}

hipError_t launch_rdet(int K, const double* B, double* rdet, hipStream_t s) {
  FESTUNG_LAUNCH(rdet_kernel, grid_for(K), dim3(BLOCK), 0, s, K, B, rdet);
  return hipGetLastError();
}

hipError_t launch_adj_stage(int p, const AdjArgs& a, hipStream_t s) {
  const int BS = a.m.bs;
  if (!valid_elem_block(BS)) return hipErrorInvalidValue;
  dim3 gr = grid_for(a.m.K, BS);
  switch (a.what) {
    case ADJ_STAGE: FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((adj_stage_kernel<PP, ADJ_STAGE>), gr, dim3(BS), 0, s, a)); break;
    case ADJ_FUSED: FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((adj_stage_kernel<PP, ADJ_FUSED>), gr, dim3(BS), 0, s, a)); break;
    case ADJ_SENS: FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((adj_stage_kernel<PP, ADJ_SENS>), gr, dim3(BS), 0, s, a)); break;
    case ADJ_SPLIT: {
      constexpr int E = 64;  // three wave-sized roles per workgroup
      dim3 g2 = grid_for(a.m.K, E);
      FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((adj_split_kernel<PP>), g2, dim3(3 * E), 0, s, a));
    } break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sens_segment(int p, const SensArgs& a, hipStream_t s) {
  if (a.count < 1 || a.first < 0 || a.nchunk < 1 || a.nchunk > a.count) return hipErrorInvalidValue;
  if (a.count % a.nchunk) return hipErrorInvalidValue;
  constexpr int E = 64;  // two wave-sized roles per workgroup
  FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((sens_seq_kernel<PP>), grid_for(a.m.K, E), dim3(2 * E), 0, s, a));
  return hipGetLastError();
}

hipError_t launch_loss(int p, int K, const double* B, const double* c, const double* cObs, const double* tab,
                       const TabOff& off, double* part, double* muT, hipStream_t s) {
  dim3 gr = grid_for(K);
  FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((loss_kernel<PP>), gr, dim3(BLOCK), 0, s, K, B, c, cObs, tab, off, part, muT));
  return hipGetLastError();
}

hipError_t launch_mass(int p, int K, const double* B, const double* x, const double* tab, const TabOff& off,
                       double* out, hipStream_t s) {
  dim3 gr = grid_for(K);
  FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((mass_kernel<PP>), gr, dim3(BLOCK), 0, s, K, B, x, tab, off, out));
  return hipGetLastError();
}

hipError_t launch_apply(int p, int op, const Mesh& m, const double* x, const double* u1, const double* u2,
                        const double* vnw, const double* tab, const TabOff& off, double* y, hipStream_t s) {
  if (op < OP_G || op > OP_MINV) return hipErrorInvalidValue;
  FESTUNG_DISPATCH_P(p, FESTUNG_LAUNCH((apply_kernel<PP>), grid_for(m.K), dim3(BLOCK), 0, s, m, op, x, u1, u2, vnw, tab,
                                       off, y));
  return hipGetLastError();
}

hipError_t launch_vertex_limiter(const Mesh& m, const double* cen, const double* valV0T, const double* dataV0T,
                                 const double* vmin, const double* vmax, const double* gmm, double* alpha,
                                 double* minV0T, double* maxV0T, hipStream_t s) {
  FESTUNG_LAUNCH(vertex_limiter_kernel, grid_for(m.K), dim3(BLOCK), 0, s, m, cen, valV0T, dataV0T, vmin, vmax, gmm,
                 alpha, minV0T, maxV0T);
  return hipGetLastError();
}

hipError_t launch_fill(size_t n, double v, double* x, hipStream_t s) {
  if (n == 0) return hipSuccess;
  FESTUNG_LAUNCH(fill_kernel, dim3(unsigned((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, n, v, x);
  return hipGetLastError();
}

int block_size() { return BLOCK; }

}  // namespace festung
