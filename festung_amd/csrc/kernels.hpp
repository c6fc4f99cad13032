// Device-side data structures shared by the kernels and the problem object.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace festung {

// Compile-time sizes for polynomial degree P.
template <int P>
struct Dims {
  static constexpr int N = (P + 1) * (P + 2) / 2;
  // points of quadRule2D(max(2P, 1)) and quadRule2D(2P + 1)
  static constexpr int RE = P == 0 ? 1 : P == 1 ? 3 : P == 2 ? 6 : P == 3 ? 12 : 64;
  static constexpr int RF = P == 0 ? 1 : P == 1 ? 4 : P == 2 ? 7 : P == 3 ? 49 : 81;
  static constexpr int Re = P + 1;  // points of quadRule1D(2P + 1)
  // points of the element term's rule (RefTables::ruleG): quadRule2D(max(2P, 1))
  // for P <= 3; at P = 4 a 28-point rule exact to degree 11
  static constexpr int RG = P == 4 ? 28 : RE;
};

// Nonzero pattern of the change to the Taylor basis (bit j of row r set when
// entry (r, j) may be nonzero), for F (DG -> Taylor: derivatives of the DG
// functions at the centroid) and its inverse.  A Taylor row of degree d only
// involves DG functions of degree >= d, and several derivatives of the
// Dubiner functions vanish at the centroid (1/3, 1/3), so 137 of the 225
// entries of F are zero at p = 4.  The kernels skip the zero products (the
// sums are unchanged: s + 0 x = s); the host checks each table against the
// pattern when it builds it.
inline constexpr uint16_t kTaylorMaskF[5][15] = {
    {0x0001},
    {0x0001, 0x0006, 0x0004},
    {0x0001, 0x003e, 0x0034, 0x0038, 0x0020, 0x0030},
    {0x0001, 0x03fe, 0x03b4, 0x03f8, 0x03a0, 0x0330, 0x03c0, 0x0380, 0x0300, 0x0200},
    {0x0001, 0x7ffe, 0x7bb4, 0x7ff8, 0x7ba0, 0x7330, 0x77c0, 0x7380, 0x7300, 0x2200, 0x7c00, 0x7800, 0x7000,
     0x6000, 0x4000}};
inline constexpr uint16_t kTaylorMaskFinv[5][15] = {
    {0x0001},
    {0x0001, 0x0006, 0x0004},
    {0x0001, 0x003e, 0x0034, 0x0038, 0x0030, 0x0010},
    {0x0001, 0x03fe, 0x03b4, 0x03f8, 0x03b0, 0x0390, 0x03c0, 0x0380, 0x0300, 0x0200},
    {0x0001, 0x7ffe, 0x7bb4, 0x7ff8, 0x7bb0, 0x7b90, 0x7fc0, 0x6380, 0x7300, 0x6200, 0x7c00, 0x7800, 0x7000,
     0x6000, 0x4000}};
__host__ __device__ constexpr bool taylor_nz(bool inv, int p, int r, int j) {
  return ((inv ? kTaylorMaskFinv[p][r] : kTaylorMaskF[p][r]) >> j) & 1;
}

// The stage kernel's element term on the matrix cores (FP64 16x16x4 MFMA)
// when the basis fits 16 rows: p = 4 (N = 15; 28 points of the element
// term's rule, in two blocks of 16 rows, the last four rows zero).  Lower
// degrees stay on the vector ALU, where the padding to 16 would cost more
// than the unit saves.
inline constexpr bool mfma_elem(int p) { return p == 4; }
template <int P>
struct MfmaElem {
  static constexpr bool value = mfma_elem(P);
  static constexpr int RG = Dims<P>::RG;
  static_assert(!value || RG % 4 == 0, "point count must fill whole 4-point steps");
  static constexpr int QB = (RG + 15) / 16;       // 16-row point blocks
  static constexpr int S = (Dims<P>::N + 3) / 4;  // 4-wide blocks of basis functions
  static constexpr int NT = RG / 4;               // 4-point steps of the second GEMM
  // 4-point steps (accumulator registers) in use in point block qb
  static constexpr int steps(int qb) { return RG - 16 * qb >= 16 ? 4 : (RG - 16 * qb) / 4; }
  // table (doubles, lane index fastest): A[qb][s][lane], D1[4qb + t][lane], D2[4qb + t][lane]
  static constexpr int offD1 = QB * S * 64, offD2 = offD1 + NT * 64, size = offD2 + NT * 64;
  // hand-over buffer of one wave: 16 rows x 64 elements
  static constexpr int xchg = 16 * 64;
};

// Offsets (in doubles) of the reference tables inside one device buffer.
struct TabOff {
  int wE, phiE;             // quadRule2D(max(2p,1)): weights, phi (RE x N)
  int wF, phiF;             // quadRule2D(2p+1)
  int wEdge, phi1D, theta;  // quadRule1D(2p+1): weights, phi1D (Re x N x 3), thetaPhi1D (Re x N x 3 x 3)
  int hatM, hatMinv;        // N x N
  // the same data laid out for the element kernels' loop order, so that the
  // values one loop iteration needs are contiguous (wide uniform loads):
  int elemQ;                // per point q of quadRule2D(max(2p,1)): w_q, phi_j(q), d1phi_j(q), d2phi_j(q)
  int edgeQ;                // per edge n, point r: w_r, phi_j(gamma_n(s_r))
  int hatMinvR, FR, FinvR;  // N x N row-major (FR, FinvR: the Taylor change of basis and its inverse)
  int FloR, FinvloR;        // FR's and FinvR's remainders: hi + lo double-double tables
  int elemG;                // elemQ's layout for the element term's rule (RefTables::ruleG)
  // operands of the MFMA element term (MfmaElem) on that rule, per lane l
  // (r = l % 16, k = l / 16): A[qb][s] = phi_(4s+k)(q = 16qb + r) (zero past
  // RG); D1/D2[4qb + t] = w_q d1phi_r(q) / w_q d2phi_r(q) at q = 16qb + 4t + k
  // (zero past N); -1 if unused
  int mfmaE;
  int total;
};

// Per-element flag word: bits 2n..2n+1 = neighbour's local edge across
// local edge n; bit 6+n = outflow (diagonal) term active on edge n; bit 9+n =
// Dirichlet inflow on edge n; bit 12+n = local vertex n lies on a Dirichlet
// edge (markV0TbdrD); bit 15+n = edge n is a Neumann edge (markE0TbdrN).
__host__ __device__ inline int flag_np(int f, int n) { return (f >> (2 * n)) & 3; }
__host__ __device__ inline bool flag_out(int f, int n) { return (f >> (6 + n)) & 1; }
__host__ __device__ inline bool flag_dir(int f, int n) { return (f >> (9 + n)) & 1; }
__host__ __device__ inline bool flag_vdir(int f, int n) { return (f >> (12 + n)) & 1; }
__host__ __device__ inline bool flag_neu(int f, int n) { return (f >> (15 + n)) & 1; }

// Coefficients of one explicit RK stage:
//   out = a0*c0 + a1*c + b*tau*cdot + g*acc
//   acc_out = d0*c0 + d1*acc + e*tau*cdot           (only when acc_out != null)
struct RKCoef {
  double a0, a1, b, g, d0, d1, e;
};

// Per-workgroup min / max of the first F Taylor coefficients a producer
// kernel writes (F = 1: the centroid values; more for the hierarchical and
// strict limiters, whose vertex bounds are taken over derivative
// coefficients too): partials[2 (f nblk + b)], partials[2 (f nblk + b) + 1].
// The vertex kernel, launched after the producer, folds them into the global
// min / max per coefficient and reads the coefficients themselves from the
// producer's Taylor output.
struct MinMaxSink {
  double* partials;
  int F;
};

struct Mesh {
  int K, numV;
  const double* B;       // 4 x K: B11, B21, B12, B22
  const double* areaE;   // 3 x K
  const int* nbr;        // 3 x K
  const int* flags;      // K
  const int* V0T;        // 3 x K (device vertex ids)
  int bs;                // elements per workgroup of the element kernels (64, 128 or 256)
  const double* rdet;    // K: 1 / (B11 B22 - B12 B21), as the kernels form it (launch_rdet)
};

// Operators of festung_apply_operator (values of FESTUNG_OP_* in the C-ABI)
enum OpKind { OP_G = 0, OP_G1 = 1, OP_G2 = 2, OP_R = 3, OP_M = 4, OP_MINV = 5 };
// OUT_DOT_DG_CEN: dc/dt in DG plus the min / max partials of its Taylor mean
// (the input of the 'linear' limiter on DG coefficients, limit_dg_kernel)
enum StageOut { OUT_RK_DG = 0, OUT_RK_TAYLOR = 1, OUT_DOT_TAYLOR = 2, OUT_DOT_DG = 3, OUT_DOT_DG_CEN = 4 };
enum AdjWhat { ADJ_STAGE = 0, ADJ_FUSED = 1, ADJ_SENS = 2, ADJ_SPLIT = 3, ADJ_SENS_U = 4, ADJ_SENS_V = 5 };
// festung_adjoint's sweep with the sensitivities deferred to one pass per
// checkpoint segment (launch_sens_segment); its stage kernels run ADJ_STAGE
constexpr int ADJ_DEFER = 6;
// LIM_SOLUTION / LIM_DOT_RK_TAYLOR: Taylor input, any limiter type.
// LIM_DOT_RK_DG / LIM_SOLUTION_DG: the 'linear' type on DG input
// (limit_dg_kernel): dc/dt limiting + RK with the new state in DG and its
// centroid partials, and the state limited in place (out == taylor).
enum LimitMode { LIM_SOLUTION = 0, LIM_DOT_RK_DG = 1, LIM_DOT_RK_TAYLOR = 2, LIM_SOLUTION_DG = 3 };
// Limiter types of core/applySlopeLimiterTaylor.m:69-85
enum LimiterType { LT_LINEAR = 0, LT_HIERARCH = 1, LT_STRICT = 2 };
// Taylor coefficients whose vertex-patch bounds a limiter type needs
__host__ __device__ inline int limiter_fields(int type, int p) { return type == LT_LINEAR ? 1 : p * (p + 1) / 2; }

struct StageArgs {
  Mesh m;
  const double* c;       // stage state (neighbours are read from it)
  const double* c0;      // state at the start of the step
  const double* acc;     // RK4 accumulator (in)
  double* accOut;        // RK4 accumulator (out) or null
  const double* u1;
  const double* u2;
  // p = 1 .. 3: 2 x RE x K, the element term's velocity factors at the points
  // q of quadRule2D(2p): w_q (B22 u1 - B12 u2)(q), w_q (B11 u2 - B21 u1)(q)
  // (launch_velocity_weights; formed whenever u1Disc or u2Disc is set)
  const double* uq;
  const double* vnw;     // Re x 3 x K: |E_kn| w_r (u . nu)(x_knr), edge_flux_weights
  const double* rhs;     // N x K or null (globL - globKD - globKN)
  const double* tab;
  TabOff off;
  double tau;
  const double* tauK;    // per-element time step (batched problems) or null
  RKCoef rk;
  double* out;           // DG or Taylor coefficients
  MinMaxSink sink;       // for the Taylor outputs
};

struct VertexArgs {
  int numV;
  int maxDeg;            // columns of the padded vertex -> element lists
  const int* v2tPad;     // maxDeg x numV, element ids (device order), -1 padded
  const double* fields;  // Taylor coefficients (N x K); the first F are reduced
  int K;
  int F;
  double scale0;         // factor on field 0: 1, or F_00 when `fields` are DG coefficients (the Taylor mean)
  const double* partials;  // producer's per-block min / max, F x nPartials x 2
  int nPartials;
  double* gmm;             // out: global min, max per coefficient (2F, written by block 0)
  double* vmin;            // out: raw patch min / max per coefficient and vertex (F x numV)
  double* vmax;
};

struct LimitArgs {
  Mesh m;
  const double* taylor;  // coefficients to limit (N x K): Taylor, or DG in the DG-input modes
  const double* vmin;    // raw patch min / max per coefficient and vertex (F x numV)
  const double* vmax;
  const double* gmm;     // global min / max per coefficient (2F)
  const double* vertexD; // Dirichlet vertex values (NaN = none) or null
  const double* tab;
  TabOff off;
  // RK combination (dc/dt modes)
  const double* c;
  const double* c0;
  const double* acc;
  double* accOut;
  double tau;
  const double* tauK;    // per-element time step or null
  RKCoef rk;
  double* out;           // DG (LIM_SOLUTION, the DG-input modes) or Taylor (LIM_DOT_RK_TAYLOR)
  MinMaxSink sink;
  double* alphaOut;      // K or null
  // LIM_DOT_RK_DG: the limited dc/dt of badly shaped elements, formed in
  // double-double by limit_dot_fixup_kernel: fixIdx (K, -1 = none) indexes
  // fixY (N x fixCount); null when the mesh has no such elements
  const int* fixIdx;
  const double* fixY;
  int fixCount;
};

struct ToTaylorArgs {
  int K;
  const double* B;
  const double* c;
  const double* tab;
  TabOff off;
  double* out;
  MinMaxSink sink;
  int bs;
};

// Vertex-based limiter of the DG coefficients' linear part,
// core/applySlopeLimiterDiscLinear.m: the vertex values of
// sum_{j<3} c_j phi_j are limited against the patch bounds of the centroid
// values c_1 phi_1(1/3, 1/3).
struct DiscLinArgs {
  Mesh m;
  const double* cen;     // centroid values (K), the field the vertex pass reduced
  const double* vmin;    // raw patch min / max of cen per vertex (numV)
  const double* vmax;
  const double* gmm;     // global min, max of cen
  const double* vertexD; // Dirichlet vertex values (NaN = none) or null
  double phiV[3][3];     // phi_(j+1) at reference vertex n: [n][j]
  double* c;             // N x K DG coefficients, limited in place
  double* alphaOut;      // K or null
};

// Adjoint stage in mu = M^-1 lambda space:
//   y = (1 - omega)(mu + tau_k M^-1 A^T mu)        (mode 0..3)
//   mode 0: out = y, L0 = omega mu;  mode 1: out = y, L0 += omega mu;
//   mode 2: out = L0 + omega mu + y; mode 3: out = omega mu + y;
//   mode 4: out = M^-1 A^T mu (no RK weights, tau ignored)
struct AdjArgs {
  Mesh m;
  const double* mu;
  const double* u1;
  const double* u2;
  const double* vn;
  const double* vnw;     // the stage kernel's edge-flux factors |E| w_r vn (read by ADJ_STAGE)
  const double* tab;
  TabOff off;
  double tau;
  const double* tauK;
  double omega;
  int mode;
  double* L0;
  double* out;
  int what;              // ADJ_STAGE, ADJ_FUSED, ADJ_SENS or ADJ_SPLIT (stage + sensitivities, three threads per element)
  // velocity sensitivities of this stage (ADJ_FUSED, ADJ_SENS):
  //   G += (1 - omega) tau_k d/d(u1Disc, u2Disc, vn) [mu^T A c]
  const double* c;       // stage state c_s
  double* Gu1;
  double* Gu2;
  double* Gvn;
};

// Velocity sensitivities of one checkpoint segment, formed after its reverse
// sweep from the adjoint stage inputs kept for it: for every element and the
// segment's adjoint stages e = first .. first + count - 1,
//   G += (1 - omega_e) tau_k d/d(u1Disc, u2Disc, vn) [mu_e^T A c_e]
// summed in nchunk equal consecutive chunks of stages (one time step each),
// the chunk sums added to G in chunk order: the same per-stage contributions
// as the stage kernels' ADJ_FUSED / ADJ_SPLIT variants, added in another
// (fixed) order.
struct SensArgs {
  Mesh m;
  const double* vn;
  const double* tab;
  TabOff off;
  double tau;
  const double* tauK;
  const double* const* mus;  // per adjoint stage: its input mu
  const double* const* cs;   // per adjoint stage: the forward stage state
  const double* omegas;      // per adjoint stage: omega of its RK stage
  int first, count, nchunk;
  double* Gu1;
  double* Gu2;
  double* Gvn;
};

hipError_t launch_stage(int p, int mode, const StageArgs& a, hipStream_t s);
// vnw = |E_kn| w_r vn (the stage kernel's edge-flux factors; the sign is vn's)
hipError_t launch_edge_flux_weights(int K, int Re, const double* areaE, const double* vn, const double* wEdge,
                                    double* vnw, hipStream_t s);
hipError_t launch_adj_stage(int p, const AdjArgs& a, hipStream_t s);
hipError_t launch_sens_segment(int p, const SensArgs& a, hipStream_t s);
// J partials: part[b] = sum over the block of det_k d_k^T hatM d_k, d = c - cObs;
// muT = 2 d
hipError_t launch_loss(int p, int K, const double* B, const double* c, const double* cObs, const double* tab,
                       const TabOff& off, double* part, double* muT, hipStream_t s);
// out = det_k hatM x_k (the block mass matrix)
hipError_t launch_mass(int p, int K, const double* B, const double* x, const double* tab, const TabOff& off,
                       double* out, hipStream_t s);
// y = A x for one reference operator applied alone (OpKind)
hipError_t launch_apply(int p, int op, const Mesh& m, const double* x, const double* u1, const double* u2,
                        const double* vnw, const double* tab, const TabOff& off, double* y, hipStream_t s);
// alpha and the shifted patch bounds of computeVertexBasedLimiter on caller data
hipError_t launch_vertex_limiter(const Mesh& m, const double* cen, const double* valV0T, const double* dataV0T,
                                 const double* vmin, const double* vmax, const double* gmm, double* alpha,
                                 double* minV0T, double* maxV0T, hipStream_t s);
hipError_t launch_fill(size_t n, double v, double* x, hipStream_t s);
hipError_t launch_vertex(const VertexArgs& a, hipStream_t s);
hipError_t launch_limit(int p, int mode, int type, const LimitArgs& a, hipStream_t s);
// the dc/dt limiter's result for the `count` listed elements (device ids) in
// double-double into y (N x count), for LIM_DOT_RK_DG's fixIdx / fixY
hipError_t launch_limit_dot_fixup(int p, const LimitArgs& a, const int* list, int count, double* y, hipStream_t s);
hipError_t launch_to_taylor(int p, const ToTaylorArgs& a, hipStream_t s);
// centroid values c_1 phi1c -> cen (K) and their min / max partials
hipError_t launch_centroid(int K, int bs, const double* c, double phi1c, double* cen, const MinMaxSink& sink,
                           hipStream_t s);
hipError_t launch_limit_disc_linear(int p, const DiscLinArgs& a, hipStream_t s);
// StageArgs::uq from u1Disc, u2Disc and B (p = 1 .. 3; nothing to do otherwise)
hipError_t launch_velocity_weights(int p, int K, const double* B, const double* u1, const double* u2, const double* tab,
                                   const TabOff& off, double* uq, hipStream_t s);
hipError_t launch_project(int p, bool ordF, int K, const double* fq, const double* tab, const TabOff& off,
                          double* out, hipStream_t s);
hipError_t launch_vnormal(int K, int Re, const double* nuE0T, const double* u1e, const double* u2e, double* vn,
                          hipStream_t s);
// rhs = globL - globKD - globKN from f at the element points (fq), c_D and
// g_N at the edge points (cDe, gNe); any of the three may be null (zero)
hipError_t launch_source(int p, const Mesh& m, const double* fq, const double* cDe, const double* gNe,
                         const double* vn, const double* tab, const TabOff& off, double* rhs, hipStream_t s);
hipError_t launch_l2(int p, bool ordF, int K, const double* B, const double* c, const double* cq, const double* tab,
                     const TabOff& off, double* partial, hipStream_t s);
hipError_t launch_permute(int K, int cols, const double* src, const int* idx, double* dst, hipStream_t s);
// rdet[k] = 1 / det(B_k) with Geo::load's operations (Mesh::rdet)
hipError_t launch_rdet(int K, const double* B, double* rdet, hipStream_t s);
int block_size();
// profiling: events recorded at the start / end of the calling thread's next
// kernels (null: plain launches)
void set_kernel_events(hipEvent_t start, hipEvent_t stop);          // workgroup size of the reduction / set-up kernels
bool valid_elem_block(int bs);  // supported elements per workgroup

}  // namespace festung
