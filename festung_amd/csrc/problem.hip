// Device-resident advection problem: buffers, the per-stage kernel
// sequence, hipGraph capture of whole steps, and the problem part of the
// C-ABI.  The stage sequence replaces advection/solveSubStep.m:57-99 with the
// operators of advection/preprocessSubStep.m:56-99 applied matrix-free.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "checkpoint.hpp"
#include "common.hpp"
#include "grid.hpp"
#include "kernels.hpp"
#include "tables.hpp"

namespace festung {

#define HIPCHK(expr)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) fail(FESTUNG_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// Owning device allocation.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t b) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    bytes = b;
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      fail(FESTUNG_ERR_ALLOC, "device allocation of %zu bytes failed: %s", b, hipGetErrorString(e));
    }
  }
  double* d() const { return static_cast<double*>(p); }
  int* i() const { return static_cast<int*>(p); }
};

struct StagePlan {
  RKCoef rk;
  int in, out;   // buffer ids: 0 = X (step start), 1 = Y, 2 = Z
  bool accOut;
};

static std::vector<StagePlan> make_plan(int ordRK) {
  std::vector<StagePlan> v;
  auto ssp = [](double w) { return RKCoef{w, 1 - w, 1 - w, 0, 0, 0, 0}; };
  switch (ordRK) {
    case 1: v = {{ssp(0.0), 0, 1, false}}; break;                                     // copy back after
    case 2: v = {{ssp(0.0), 0, 1, false}, {ssp(0.5), 1, 0, false}}; break;
    case 3: v = {{ssp(0.0), 0, 1, false}, {ssp(3.0 / 4), 1, 2, false}, {ssp(1.0 / 3), 2, 0, false}}; break;
    case 4:  // classic RK4 in accumulator form (extension)
      v = {{RKCoef{1, 0, 0.5, 0, 1, 0, 1.0 / 6}, 0, 1, true},
           {RKCoef{1, 0, 0.5, 0, 0, 1, 1.0 / 3}, 1, 2, true},
           {RKCoef{1, 0, 1.0, 0, 0, 1, 1.0 / 3}, 2, 1, true},
           {RKCoef{0, 0, 1.0 / 6, 1, 0, 0, 0}, 1, 0, false}};
      break;
    default: fail(FESTUNG_ERR_UNSUPPORTED, "Runge-Kutta order %d not supported (1..4)", ordRK);
  }
  return v;
}

// Hilbert-curve index of (x, y) in [0, n)^2, n a power of two (the
// classic iterative quadrant-rotation construction).
static uint64_t hilbert_index(uint32_t n, uint32_t x, uint32_t y) {
  uint64_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    const uint32_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
    d += uint64_t(s) * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

// Order of points along a Hilbert curve over their bounding box: perm[d] is
// the id placed at position d.  Ties (same 2^16 grid cell) keep id order.
static std::vector<int32_t> hilbert_order(const double* x, const double* y, int32_t n) {
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (int32_t i = 0; i < n; ++i) {
    x0 = std::min(x0, x[i]); x1 = std::max(x1, x[i]);
    y0 = std::min(y0, y[i]); y1 = std::max(y1, y[i]);
  }
  const double span = std::max(std::max(x1 - x0, y1 - y0), 1e-300);
  const uint32_t G = 1u << 16;
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  for (int32_t i = 0; i < n; ++i) {
    uint32_t qx = uint32_t(std::min(double(G - 1), std::floor((x[i] - x0) / span * G)));
    uint32_t qy = uint32_t(std::min(double(G - 1), std::floor((y[i] - y0) / span * G)));
    key[i] = {hilbert_index(G, qx, qy), i};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> perm(n);
  for (int32_t i = 0; i < n; ++i) perm[i] = key[i].second;
  return perm;
}

}  // namespace festung

using namespace festung;

namespace {
// Tuning-key device part: architecture name and compute-unit count.
std::string device_tag(int device) {
  static std::mutex mu;
  static std::vector<std::string> tags;
  std::lock_guard<std::mutex> lk(mu);
  if (device >= int(tags.size())) tags.resize(device + 1);
  if (tags[device].empty()) {
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    std::string arch(prop.gcnArchName);
    for (char& ch : arch)
      if (ch == ' ') ch = '_';
    tags[device] = arch + "/" + std::to_string(prop.multiProcessorCount) + "cu";
  }
  return tags[device];
}

int ceil_log2(int K) {
  int l = 0;
  while ((int64_t(1) << l) < K) ++l;
  return l;
}
}  // namespace

struct festung_problem {
  int device = 0;
  hipStream_t stream = nullptr;
  int p = 0, N = 0, K = 0, numV = 0, ordRK = 0, flags = 0;
  RefTables tabs;
  TabOff off{};
  std::vector<double> areaT;  // host copy (L2 error)
  // Device element / vertex order: Hilbert-curve order of the centroids /
  // vertices, so that neighbours and vertex patches are close in memory.
  // eperm[d] = reference element at device position d, einv = inverse;
  // vperm / vinv the same for vertices.  All C-ABI arrays stay in the
  // reference's order and are permuted here.
  std::vector<int32_t> eperm, einv, vperm, vinv, V0Tref;
  DBuf dEperm, dEinv, dVperm;
  DBuf devQ1, devQ2, devE1, devE2;  // device-order copies of point values given by device pointer
  std::vector<StagePlan> plan;
  DBuf tab, B, rdet, area, nbr, flg, V0T, v2tPad, nuE0T;
  int maxDeg = 0;
  DBuf tauK;
  bool hasTauK = false;
  DBuf X, Y, Z, acc, T1, T2, partials, gmm, vmin, vmax, vertexD, u1, u2, vn, rhs, alpha;
  DBuf vnw;  // the stage kernel's edge-flux factors |E| w_r vn, formed whenever vn is set
  DBuf uq;   // the stage kernel's velocity factors at the element points (p = 1 .. 3), formed with u1, u2
  int limType = LT_LINEAR;  // limiter type (core/applySlopeLimiterTaylor.m:69-85)
  int nF = 1;               // Taylor coefficients with vertex-patch bounds (limiter_fields)
  // badly shaped elements (cond(B)^p > 4096), whose limited dc/dt the
  // double-double fix-up kernel forms ('linear' limiter, p >= 2): device ids,
  // per-element index into fixY (-1: none), fixY N x fixCount
  DBuf fixList, fixIdx, fixY;
  int fixCount = 0;
  bool hasRhs = false, hasVertexD = false, hasVelocity = false;
  int velocityParts = 0;  // fields set by festung_set_velocity_disc: 1 u1Disc, 2 u2Disc, 4 vNormal
  int cur = 0;           // buffer holding the current state (0 = X)
  int nextStage = -1;    // next sub-step of a step begun with festung_begin_step
  // captured step sequences: graph[0] one step, graph[1] kGraphSteps steps
  // (fewer graph launches: each boundary between launches costs a few us)
  static constexpr int kGraphSteps = 8;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t graphExec[2] = {nullptr, nullptr};
  double graphTau = std::numeric_limits<double>::quiet_NaN();
  int bs = 256;      // elements per workgroup of the element kernels
  int nblocks = 0;   // workgroups of the element kernels = min/max partials
  int nred = 0;      // workgroups of the reduction kernels (L2 error, misfit)
  int version = 0;   // bumped when a change invalidates captured step sequences
  TuneKey tune_key() const {
    return TuneKey{device_tag(device), p, ordRK, (flags & FESTUNG_FLAG_LIMITER) ? limType : -1, ceil_log2(K)};
  }
  // per-launch timing (HIP events on the problem's stream)
  bool prof = false;
  struct Rec {
    int kind;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;

  hipEvent_t get_event() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  template <class F>
  void timed(int kind, F&& f) {
    if (!prof) {
      f();
      return;
    }
    Rec r{kind, get_event(), get_event()};
    if (kind == FESTUNG_KERNEL_COPY) {
      HIPCHK(hipEventRecord(r.a, stream));
      f();
      HIPCHK(hipEventRecord(r.b, stream));
    } else {
      // the kernel's own start / end (not the launch's)
      struct Reset {
        ~Reset() { set_kernel_events(nullptr, nullptr); }
      } reset;
      set_kernel_events(r.a, r.b);
      f();
    }
    recs.push_back(r);
  }

  double* buf(int id) const { return id == 0 ? X.d() : id == 1 ? Y.d() : Z.d(); }
  Mesh mesh() const {
    return Mesh{K, numV, B.d(), area.d(), nbr.i(), flg.i(), V0T.i(), bs, rdet.d()};
  }
  MinMaxSink sink() const {
    return MinMaxSink{partials.d(), nF};
  }
  void set_device() const { HIPCHK(hipSetDevice(device)); }
  void drop_graph() {
    for (int i = 0; i < 2; ++i) {
      if (graphExec[i]) (void)hipGraphExecDestroy(graphExec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
      graphExec[i] = nullptr;
      graph[i] = nullptr;
    }
    graphTau = std::numeric_limits<double>::quiet_NaN();
  }
  hipStream_t aux[2] = {nullptr, nullptr};  // side streams of the adjoint sweep
  void ensure_aux_streams() {
    for (auto& a : aux)
      if (!a) HIPCHK(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
  }
  ~festung_problem() {
    drop_graph();
    for (auto a : aux)
      if (a) (void)hipStreamDestroy(a);
    for (auto& r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <class T>
  void upload(DBuf& b, const T* h, size_t n) {
    HIPCHK(hipStreamSynchronize(stream));  // kernels in flight may still use b
    b.alloc(n * sizeof(T));
    HIPCHK(hipMemcpy(b.p, h, n * sizeof(T), hipMemcpyHostToDevice));
  }
  // K x cols column-major array, reference order -> device order
  std::vector<double> to_dev(const double* src, int cols) const {
    std::vector<double> out(size_t(K) * cols);
    for (int j = 0; j < cols; ++j)
      for (int d = 0; d < K; ++d) out[d + size_t(j) * K] = src[eperm[d] + size_t(j) * K];
    return out;
  }
  void from_dev(const double* dev, int cols, double* dst) const {
    for (int j = 0; j < cols; ++j)
      for (int d = 0; d < K; ++d) dst[eperm[d] + size_t(j) * K] = dev[d + size_t(j) * K];
  }
  void upload_elem(DBuf& b, const double* src, int cols, bool realloc = true) {
    std::vector<double> v = to_dev(src, cols);
    // the copy goes through the null stream, which does not wait for the
    // problem's (non-blocking) stream: let kernels in flight finish first
    HIPCHK(hipStreamSynchronize(stream));
    if (realloc || !b.p) b.alloc(v.size() * sizeof(double));
    HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  void download_elem(const DBuf& b, const double* devp, int cols, double* dst) {
    (void)b;
    std::vector<double> v(size_t(K) * cols);
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(v.data(), devp, v.size() * sizeof(double), hipMemcpyDeviceToHost));
    from_dev(v.data(), cols, dst);
  }
  // Caller arrays of the entry point running now: host memory (reference
  // order, copied) or, while ioDev is set (the _dev entry points), device
  // memory in the same layout, permuted on the device.
  bool ioDev = false;
  void load_in(DBuf& b, const double* src, int cols) {
    if (!ioDev) {
      upload_elem(b, src, cols);
      return;
    }
    b.alloc(size_t(K) * cols * sizeof(double));
    HIPCHK(launch_permute(K, cols, src, dEperm.i(), b.d(), stream));
  }
  void store_out(const DBuf& b, int cols, double* dst) {
    if (!ioDev) {
      download_elem(b, b.d(), cols, dst);
      return;
    }
    HIPCHK(launch_permute(K, cols, b.d(), dEinv.i(), dst, stream));
    HIPCHK(hipStreamSynchronize(stream));  // b may be freed when the entry point returns
  }

  // raw vertex-patch min / max of the first nF Taylor coefficients just
  // written to `fields` (the centroid values first), and their global min /
  // max from the producer's per-block partials
  // Captured step sequences bake in whether Dirichlet vertex data are used:
  // a change invalidates them.
  void set_has_vertexD(bool v) {
    if (v == hasVertexD) return;
    HIPCHK(hipStreamSynchronize(stream));
    drop_graph();
    ++version;
    hasVertexD = v;
  }
  // The same for whether a stage source (globL - globKD - globKN) is added.
  void set_has_rhs(bool v) {
    if (v == hasRhs) return;
    HIPCHK(hipStreamSynchronize(stream));
    drop_graph();
    ++version;
    hasRhs = v;
  }
  // Captured sequences also bake in the scalar time step: drop them (after
  // the launches in flight, which may still run them) when it changes.
  void drop_graph_synced() {
    HIPCHK(hipStreamSynchronize(stream));
    drop_graph();
  }

  void refresh_uq() {
    HIPCHK(launch_velocity_weights(p, K, B.d(), u1.d(), u2.d(), tab.d(), off, uq.d(), stream));
  }
  void refresh_vnw() {
    HIPCHK(launch_edge_flux_weights(K, tabs.Re, area.d(), vn.d(), tab.d() + off.wEdge, vnw.d(), stream));
  }

  void enqueue_vertex(const double* fields, int F = -1, double scale0 = 1.0) {
    VertexArgs va{numV,   maxDeg, v2tPad.i(), fields,   K,        F < 0 ? nF : F,
                  scale0, partials.d(), nblocks, gmm.d(), vmin.d(), vmax.d()};
    timed(FESTUNG_KERNEL_VERTEX, [&] { HIPCHK(launch_vertex(va, stream)); });
  }

  // One RK stage: state buffer `in` -> `out`, step start in X.
  // keepAlpha: write the solution limiter's alpha (read by festung_limiter_info;
  // only the last stage of a whole step needs to)
  // stage-kernel arguments of one sub-step (output pointer set by the caller)
  StageArgs stage_args(const StagePlan& sp, double tau) {
    StageArgs a{};
    a.m = mesh();
    a.c = buf(sp.in);
    a.c0 = X.d();
    a.acc = acc.d();
    a.accOut = sp.accOut ? acc.d() : nullptr;
    a.u1 = u1.d();
    a.u2 = u2.d();
    a.uq = uq.d();
    a.vnw = vnw.d();
    a.rhs = hasRhs ? rhs.d() : nullptr;
    a.tab = tab.d();
    a.off = off;
    a.tau = tau;
    a.tauK = hasTauK ? tauK.d() : nullptr;
    a.rk = sp.rk;
    a.sink = sink();
    return a;
  }

  void enqueue_stage(const StagePlan& sp, double tau, bool keepAlpha = true) {
    const bool lim = (flags & FESTUNG_FLAG_LIMITER) != 0;
    const double* cin = buf(sp.in);
    double* cout = buf(sp.out);
    StageArgs a = stage_args(sp, tau);
    if (!lim) {
      a.out = cout;
      timed(FESTUNG_KERNEL_STAGE, [&] { HIPCHK(launch_stage(p, OUT_RK_DG, a, stream)); });
      return;
    }
    LimitArgs l{};
    l.m = mesh();
    l.vmin = vmin.d();
    l.vmax = vmax.d();
    l.gmm = gmm.d();
    l.vertexD = nullptr;
    l.tab = tab.d();
    l.off = off;
    l.sink = sink();
    l.alphaOut = alpha.d();
    if (p >= 2 && limType == LT_LINEAR) {
      // the same three steps on DG coefficients (limit_dg_kernel): dc/dt in
      // T1, the new state written to `out` and limited there in place; the
      // vertex passes reduce the Taylor means F_00 c_1
      const double f00 = tabs.F[0];
      a.out = T1.d();
      timed(FESTUNG_KERNEL_STAGE, [&] { HIPCHK(launch_stage(p, OUT_DOT_DG_CEN, a, stream)); });
      enqueue_vertex(T1.d(), -1, f00);
      l.taylor = T1.d();
      if (fixCount > 0) {
        HIPCHK(launch_limit_dot_fixup(p, l, fixList.i(), fixCount, fixY.d(), stream));
        l.fixIdx = fixIdx.i();
        l.fixY = fixY.d();
        l.fixCount = fixCount;
      }
      l.c = cin;
      l.c0 = X.d();
      l.acc = acc.d();
      l.accOut = sp.accOut ? acc.d() : nullptr;
      l.tau = tau;
      l.tauK = hasTauK ? tauK.d() : nullptr;
      l.rk = sp.rk;
      l.out = cout;
      l.alphaOut = nullptr;
      timed(FESTUNG_KERNEL_LIMIT_DOT, [&] { HIPCHK(launch_limit(p, LIM_DOT_RK_DG, limType, l, stream)); });
      enqueue_vertex(cout, -1, f00);
      l.taylor = cout;
      l.vertexD = hasVertexD ? vertexD.d() : nullptr;
      l.alphaOut = keepAlpha ? alpha.d() : nullptr;
      timed(FESTUNG_KERNEL_LIMIT, [&] { HIPCHK(launch_limit(p, LIM_SOLUTION_DG, limType, l, stream)); });
      return;
    }
    if (p >= 2) {
      // dc/dt limiting (advection/solveSubStep.m:69-79), then RK, then
      // solution limiting (:86-97)
      a.out = T1.d();
      timed(FESTUNG_KERNEL_STAGE, [&] { HIPCHK(launch_stage(p, OUT_DOT_TAYLOR, a, stream)); });
      enqueue_vertex(T1.d());
      l.taylor = T1.d();
      l.c = cin;
      l.c0 = X.d();
      l.acc = acc.d();
      l.accOut = sp.accOut ? acc.d() : nullptr;
      l.tau = tau;
      l.tauK = hasTauK ? tauK.d() : nullptr;
      l.rk = sp.rk;
      l.out = T2.d();
      l.alphaOut = nullptr;
      timed(FESTUNG_KERNEL_LIMIT_DOT, [&] { HIPCHK(launch_limit(p, LIM_DOT_RK_TAYLOR, limType, l, stream)); });
    } else {
      a.out = T2.d();
      timed(FESTUNG_KERNEL_STAGE, [&] { HIPCHK(launch_stage(p, OUT_RK_TAYLOR, a, stream)); });
    }
    enqueue_vertex(T2.d());
    l.taylor = T2.d();
    l.out = cout;
    l.vertexD = hasVertexD ? vertexD.d() : nullptr;
    l.alphaOut = keepAlpha ? alpha.d() : nullptr;
    timed(FESTUNG_KERNEL_LIMIT, [&] { HIPCHK(launch_limit(p, LIM_SOLUTION, limType, l, stream)); });
  }

  void enqueue_step(double tau) {
    for (size_t i = 0; i < plan.size(); ++i) enqueue_stage(plan[i], tau, i + 1 == plan.size());
    if (plan.back().out != 0)
      timed(FESTUNG_KERNEL_COPY, [&] {
        HIPCHK(hipMemcpyAsync(X.p, buf(plan.back().out), size_t(N) * K * sizeof(double), hipMemcpyDeviceToDevice,
                              stream));
      });
  }

  void require_velocity() const {
    if (!hasVelocity) fail(FESTUNG_ERR_STATE, "velocity not set (festung_set_velocity_disc / _points)");
  }

  void finish_stage_sequence() {
    if (nextStage >= 0) fail(FESTUNG_ERR_STATE, "a step begun with festung_begin_step is not finished");
  }
};

namespace {

void check_handle(const festung_problem* h) {
  if (!h) fail(FESTUNG_ERR_ARG, "null problem handle");
}

void sync(festung_problem* h) { HIPCHK(hipStreamSynchronize(h->stream)); }

}  // namespace

extern "C" {

int festung_problem_create(const festung_grid* gh, int p, int ordRK, int flags, festung_problem** out) {
  return guard([&] {
    if (!gh || !out) fail(FESTUNG_ERR_ARG, "null pointer argument");
    *out = nullptr;
    if (p < 0 || p > 4) fail(FESTUNG_ERR_UNSUPPORTED, "polynomial degree p = %d (must be 0..4)", p);
    if ((flags & FESTUNG_FLAG_LIMITER) && p < 1)
      fail(FESTUNG_ERR_UNSUPPORTED, "slope limiting needs p >= 1 (advection/configureProblem.m:103)");
    const Grid& g = gh->g;
    std::unique_ptr<festung_problem> h(new festung_problem);
    h->plan = make_plan(ordRK);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) fail(FESTUNG_ERR_HIP, "no HIP device");
    HIPCHK(hipGetDevice(&h->device));
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->p = p;
    h->ordRK = ordRK;
    h->flags = flags;
    h->tabs = build_ref_tables(p);
    const int N = h->N = h->tabs.N;
    const int K = h->K = g.numT;
    h->numV = g.numV;
    h->areaT = g.areaT;  // reference order
    if (int64_t(K) * N > (int64_t(1) << 31) - 1) fail(FESTUNG_ERR_ARG, "K*N exceeds int32 indexing");

    // reference tables, one buffer
    std::vector<double> tb;
    auto put = [&](const std::vector<double>& v) {
      int o = int(tb.size());
      tb.insert(tb.end(), v.begin(), v.end());
      return o;
    };
    const RefTables& t = h->tabs;
    TabOff& o = h->off;
    o.wE = put(t.ruleE.w);
    o.phiE = put(t.phiE);
    o.wF = put(t.ruleF.w);
    o.phiF = put(t.phiF);
    o.wEdge = put(t.ruleEdge.w);
    o.phi1D = put(t.phi1D);
    o.theta = put(t.thetaPhi1D);
    o.hatM = put(t.hatM);
    o.hatMinv = put(t.hatMinv);
    {
      // loop-ordered copies (see TabOff)
      const int RE = t.RE, Re = t.Re;
      std::vector<double> eq, ed, mi(size_t(N) * N), fr(size_t(N) * N), fir(size_t(N) * N);
      std::vector<double> flo(size_t(N) * N), filo(size_t(N) * N);
      for (int q = 0; q < RE; ++q) {
        eq.push_back(t.ruleE.w[q]);
        for (int j = 0; j < N; ++j) eq.push_back(t.phiE[q + j * RE]);
        for (int m = 0; m < 2; ++m)
          for (int j = 0; j < N; ++j) eq.push_back(t.gradE[q + j * RE + m * RE * N]);
      }
      for (int n = 0; n < 3; ++n)
        for (int r = 0; r < Re; ++r) {
          ed.push_back(t.ruleEdge.w[r]);
          for (int j = 0; j < N; ++j) ed.push_back(t.phi1D[r + j * Re + n * Re * N]);
        }
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          mi[size_t(i) * N + j] = t.hatMinv[i + size_t(j) * N];
          fr[size_t(i) * N + j] = t.F[i + size_t(j) * N];
          fir[size_t(i) * N + j] = t.Finv[i + size_t(j) * N];
          flo[size_t(i) * N + j] = t.Flo[i + size_t(j) * N];
          filo[size_t(i) * N + j] = t.Finvlo[i + size_t(j) * N];
          // the kernels skip the products outside the compile-time pattern
          if ((fr[size_t(i) * N + j] != 0 && !taylor_nz(false, p, i, j)) ||
              (fir[size_t(i) * N + j] != 0 && !taylor_nz(true, p, i, j)))
            fail(FESTUNG_ERR_INTERNAL, "Taylor change of basis (%d, %d) outside its nonzero pattern", i, j);
        }
      o.elemQ = put(eq);
      if (t.RG == t.RE) {
        o.elemG = o.elemQ;
      } else {
        std::vector<double> eg;
        for (int q = 0; q < t.RG; ++q) {
          eg.push_back(t.ruleG.w[q]);
          for (int j = 0; j < N; ++j) eg.push_back(t.phiG[q + j * t.RG]);
          for (int m = 0; m < 2; ++m)
            for (int j = 0; j < N; ++j) eg.push_back(t.gradG[q + j * t.RG + m * t.RG * N]);
        }
        o.elemG = put(eg);
      }
      o.edgeQ = put(ed);
      o.hatMinvR = put(mi);
      o.FR = put(fr);
      o.FinvR = put(fir);
      o.FloR = put(flo);
      o.FinvloR = put(filo);
      if (p == 4) {
        // the kernels' mass solve takes hatM as the identity at p = 4
        for (int i = 0; i < N; ++i)
          for (int j = 0; j < N; ++j)
            if (std::fabs(t.hatM[i + size_t(j) * N] - (i == j ? 1.0 : 0.0)) > 1e-13)
              fail(FESTUNG_ERR_INTERNAL, "hatM(%d, %d) off the identity", i, j);
      }
      o.mfmaE = -1;
      if (mfma_elem(p)) {
        // lane-distributed operands of the MFMA element term (see TabOff::mfmaE)
        const int RG = t.RG, QBn = (RG + 15) / 16, S = (N + 3) / 4, NT = RG / 4;
        if (RG % 4 != 0 || RG != Dims<4>::RG) fail(FESTUNG_ERR_INTERNAL, "element-term rule of %d points", RG);
        std::vector<double> mt(size_t(QBn * S + 2 * NT) * 64, 0.0);
        for (int qb = 0; qb < QBn; ++qb)
          for (int l = 0; l < 64; ++l) {
            const int r = l % 16, k = l / 16;
            for (int s = 0; s < S; ++s) {
              const int q = 16 * qb + r, j = 4 * s + k;
              if (j < N && q < RG) mt[(size_t(qb) * S + s) * 64 + l] = t.phiG[q + size_t(j) * RG];
            }
            for (int u = 0; u < 4 && 16 * qb + 4 * u < RG; ++u) {
              const int q = 16 * qb + 4 * u + k, i = r;
              if (i >= N) continue;
              const size_t d1 = size_t(QBn) * S * 64 + (size_t(qb) * 4 + u) * 64 + l;
              const size_t d2 = d1 + size_t(NT) * 64;
              mt[d1] = t.ruleG.w[q] * t.gradG[q + size_t(i) * RG];
              mt[d2] = t.ruleG.w[q] * t.gradG[q + size_t(i) * RG + size_t(RG) * N];
            }
          }
        if (mt.size() != size_t(MfmaElem<4>::size)) fail(FESTUNG_ERR_INTERNAL, "MFMA element table size");
        o.mfmaE = put(mt);
      }
    }
    o.total = int(tb.size());
    h->upload(h->tab, tb.data(), tb.size());

    // device order of elements and vertices (Hilbert curve)
    // (a union of grids: each part along its own curve, parts in order, so
    // that a part's elements, and its neighbour gathers, stay together --
    // the parts all cover the same square, and one curve over the union
    // interleaved them)
    auto order = [](const std::vector<int32_t>& part, const double* xy, int32_t n) {
      if (part.size() < 3) return hilbert_order(xy, xy + n, n);
      std::vector<int32_t> perm;
      perm.reserve(n);
      for (size_t i = 0; i + 1 < part.size(); ++i) {
        const int32_t a = part[i], b = part[i + 1];
        for (int32_t x : hilbert_order(xy + a, xy + n + a, b - a)) perm.push_back(x + a);
      }
      return perm;
    };
    h->eperm = order(g.partT, g.baryT.data(), K);
    h->vperm = order(g.partV, g.coordV.data(), g.numV);
    h->einv.assign(K, 0);
    h->vinv.assign(g.numV, 0);
    for (int32_t d = 0; d < K; ++d) h->einv[h->eperm[d]] = d;
    for (int32_t d = 0; d < g.numV; ++d) h->vinv[h->vperm[d]] = d;
    h->V0Tref = g.V0T;
    h->upload(h->dEperm, h->eperm.data(), h->eperm.size());
    h->upload(h->dVperm, h->vperm.data(), h->vperm.size());
    h->upload(h->dEinv, h->einv.data(), h->einv.size());

    // mesh, in device order
    h->upload_elem(h->B, g.B.data(), 4);
    h->rdet.alloc(size_t(h->K) * sizeof(double));
    HIPCHK(launch_rdet(h->K, h->B.d(), h->rdet.d(), h->stream));
    h->upload_elem(h->area, g.areaE0T.data(), 3);
    h->upload_elem(h->nuE0T, g.nuE0T.data(), 6);
    {
      std::vector<int32_t> v0t(3 * size_t(K));
      for (int n = 0; n < 3; ++n)
        for (int32_t d = 0; d < K; ++d) v0t[d + size_t(n) * K] = h->vinv[g.V0T[h->eperm[d] + size_t(n) * K]];
      h->upload(h->V0T, v0t.data(), v0t.size());
      // padded vertex -> element lists, maxDeg x numV (device orders)
      int md = 0;
      for (int32_t v = 0; v < g.numV; ++v) md = std::max(md, g.v2tPtr[v + 1] - g.v2tPtr[v]);
      h->maxDeg = md;
      std::vector<int32_t> pad(size_t(md) * g.numV, -1);
      for (int32_t dv = 0; dv < g.numV; ++dv) {
        const int32_t v = h->vperm[dv];
        for (int32_t i = g.v2tPtr[v], j = 0; i < g.v2tPtr[v + 1]; ++i, ++j)
          pad[size_t(j) * g.numV + dv] = h->einv[g.v2tIdx[i]];
      }
      h->upload(h->v2tPad, pad.data(), pad.size());
    }
    // neighbours and flags (markE0TE0T, markElemEdges, markE0TbdrN/D of
    // core/assembleMatEdgePhiPhiValUpwind.m:261-299)
    std::vector<int32_t> nb(3 * size_t(K)), fl(K, 0);
    for (int32_t d = 0; d < K; ++d)
      for (int n = 0; n < 3; ++n) {
        const int32_t k = h->eperm[d];
        const size_t kn = k + size_t(n) * K, dn = d + size_t(n) * K;
        const bool notN = !g.markE0TbdrN[kn];
        const bool hasNb = g.nbr[kn] >= 0;
        nb[dn] = (hasNb && notN) ? h->einv[g.nbr[kn]] : -1;
        const bool markElemEdges = hasNb || g.idE0T[kn] != 0;
        if (hasNb) fl[d] |= (g.nbrEdge[kn] & 3) << (2 * n);
        if (notN && markElemEdges) fl[d] |= 1 << (6 + n);
        if (g.markE0TbdrD[kn]) fl[d] |= 1 << (9 + n);
        if (g.markV0TbdrD[kn]) fl[d] |= 1 << (12 + n);
        if (g.markE0TbdrN[kn]) fl[d] |= 1 << (15 + n);
      }
    h->upload(h->nbr, nb.data(), nb.size());
    h->upload(h->flg, fl.data(), fl.size());

    // badly shaped elements: cond(B)^p above 4096, where the closed-form
    // change to the Taylor basis and back can lose more than about 1e-12
    // (none on the structured meshes: cond(B) <= 2.62 there)
    if (p >= 2) {
      std::vector<int32_t> list, idx(K, -1);
      for (int32_t d = 0; d < K; ++d) {
        const int32_t k = h->eperm[d];
        const double a = g.B[k], c = g.B[k + size_t(K)], b = g.B[k + 2 * size_t(K)], e = g.B[k + 3 * size_t(K)];
        const double S = a * a + b * b + c * c + e * e, det = std::fabs(a * e - b * c);
        const double r = std::sqrt(std::max(S * S - 4.0 * det * det, 0.0));
        const double smax = std::sqrt(0.5 * (S + r)), smin = det / smax;  // singular values of B
        const double cond = smin > 0.0 ? smax / smin : INFINITY;
        if (std::pow(cond, p) > 4096.0) {
          idx[d] = int32_t(list.size());
          list.push_back(d);
        }
      }
      h->fixCount = int(list.size());
      if (h->fixCount > 0) {
        h->upload(h->fixList, list.data(), list.size());
        h->upload(h->fixIdx, idx.data(), idx.size());
        h->fixY.alloc(size_t(N) * h->fixCount * sizeof(double));
      }
    }

    // state and work buffers
    const size_t nk = size_t(N) * K * sizeof(double);
    h->X.alloc(nk);
    h->Y.alloc(nk);
    h->Z.alloc(nk);
    if (ordRK == 4) h->acc.alloc(nk);
    h->u1.alloc(nk);
    h->u2.alloc(nk);
    h->vn.alloc(size_t(3) * t.Re * K * sizeof(double));
    h->vnw.alloc(size_t(3) * t.Re * K * sizeof(double));
    if (p >= 1 && p <= 3) h->uq.alloc(size_t(2) * t.RE * K * sizeof(double));
    HIPCHK(hipMemset(h->X.p, 0, nk));
    // default elements per workgroup: the largest size that still gives at
    // least 1024 workgroups (4 per compute unit), 64 below that
    if ((flags & FESTUNG_FLAG_LIM_HIERARCH) && (flags & FESTUNG_FLAG_LIM_STRICT))
      fail(FESTUNG_ERR_ARG, "choose one limiter type");
    h->limType = (flags & FESTUNG_FLAG_LIM_HIERARCH) ? LT_HIERARCH : (flags & FESTUNG_FLAG_LIM_STRICT) ? LT_STRICT : LT_LINEAR;
    h->nF = limiter_fields(h->limType, p);
    h->bs = K >= 1024 * 256 ? 256 : K >= 1024 * 128 ? 128 : 64;
    // an autotuner result for this kind of problem wins over the default
    tune_load_env();
    if (const int tuned = tune_lookup(h->tune_key())) h->bs = tuned;
    h->nblocks = (K + h->bs - 1) / h->bs;
    h->nred = (K + block_size() - 1) / block_size();
    if (flags & FESTUNG_FLAG_LIMITER) {
      h->T1.alloc(nk);
      h->T2.alloc(nk);
      h->alpha.alloc(size_t(K) * sizeof(double));
      h->vmin.alloc(size_t(h->nF) * g.numV * sizeof(double));
      h->vmax.alloc(size_t(h->nF) * g.numV * sizeof(double));
      h->vertexD.alloc(size_t(g.numV) * sizeof(double));
    }
    // room for 64 elements per workgroup
    h->partials.alloc(size_t(2) * h->nF * ((K + 63) / 64 + 1) * sizeof(double));
    h->gmm.alloc(2 * size_t(h->nF) * sizeof(double));
    HIPCHK(hipDeviceSynchronize());
    *out = h.release();
  });
}

void festung_problem_destroy(festung_problem* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
}

int festung_problem_sizes(const festung_problem* h, int32_t* K, int32_t* N, int32_t* RE, int32_t* RF, int32_t* Re) {
  return guard([&] {
    check_handle(h);
    if (K) *K = h->K;
    if (N) *N = h->N;
    if (RE) *RE = h->tabs.RE;
    if (RF) *RF = h->tabs.RF;
    if (Re) *Re = h->tabs.Re;
  });
}

int festung_problem_stream(const festung_problem* h, void** stream) {
  return guard([&] {
    check_handle(h);
    if (!stream) fail(FESTUNG_ERR_ARG, "null pointer argument");
    *stream = h->stream;
  });
}

int festung_set_velocity_disc(festung_problem* h, const double* u1Disc, const double* u2Disc, const double* vNormal) {
  return guard([&] {
    check_handle(h);
    if (!u1Disc && !u2Disc && !vNormal) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    if (u1Disc) h->upload_elem(h->u1, u1Disc, h->N, false);
    if (u2Disc) h->upload_elem(h->u2, u2Disc, h->N, false);
    if (u1Disc || u2Disc) h->refresh_uq();
    if (vNormal) {
      h->upload_elem(h->vn, vNormal, 3 * h->tabs.Re, false);
      h->refresh_vnw();
    }
    h->velocityParts |= (u1Disc ? 1 : 0) | (u2Disc ? 2 : 0) | (vNormal ? 4 : 0);
    h->hasVelocity = h->velocityParts == 7;
  });
}

int festung_set_velocity_points(festung_problem* h, const double* u1q, const double* u2q, const double* u1e,
                                const double* u2e) {
  return guard([&] {
    check_handle(h);
    if (!u1q || !u2q || !u1e || !u2e) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    const int K = h->K;
    DBuf q, e1, e2;
    h->upload_elem(q, u1q, h->tabs.RE);
    HIPCHK(launch_project(h->p, false, K, q.d(), h->tab.d(), h->off, h->u1.d(), h->stream));
    sync(h);
    h->upload_elem(q, u2q, h->tabs.RE, false);
    HIPCHK(launch_project(h->p, false, K, q.d(), h->tab.d(), h->off, h->u2.d(), h->stream));
    h->refresh_uq();
    h->upload_elem(e1, u1e, 3 * h->tabs.Re);
    h->upload_elem(e2, u2e, 3 * h->tabs.Re);
    HIPCHK(launch_vnormal(K, h->tabs.Re, h->nuE0T.d(), e1.d(), e2.d(), h->vn.d(), h->stream));
    h->refresh_vnw();
    sync(h);
    h->velocityParts = 7;
    h->hasVelocity = true;
  });
}

int festung_set_velocity_points_dev(festung_problem* h, const double* u1q, const double* u2q, const double* u1e,
                                    const double* u2e) {
  return guard([&] {
    check_handle(h);
    if (!u1q || !u2q || !u1e || !u2e) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    const int K = h->K, RE = h->tabs.RE, Re3 = 3 * h->tabs.Re;
    if (!h->devQ1.p) h->devQ1.alloc(size_t(RE) * K * sizeof(double));
    if (!h->devE1.p) h->devE1.alloc(size_t(Re3) * K * sizeof(double));
    if (!h->devE2.p) h->devE2.alloc(size_t(Re3) * K * sizeof(double));
    // kernels still in flight read u1, u2, vn: the stream orders the update after them
    HIPCHK(launch_permute(K, RE, u1q, h->dEperm.i(), h->devQ1.d(), h->stream));
    HIPCHK(launch_project(h->p, false, K, h->devQ1.d(), h->tab.d(), h->off, h->u1.d(), h->stream));
    HIPCHK(launch_permute(K, RE, u2q, h->dEperm.i(), h->devQ1.d(), h->stream));
    HIPCHK(launch_project(h->p, false, K, h->devQ1.d(), h->tab.d(), h->off, h->u2.d(), h->stream));
    h->refresh_uq();
    HIPCHK(launch_permute(K, Re3, u1e, h->dEperm.i(), h->devE1.d(), h->stream));
    HIPCHK(launch_permute(K, Re3, u2e, h->dEperm.i(), h->devE2.d(), h->stream));
    HIPCHK(launch_vnormal(K, h->tabs.Re, h->nuE0T.d(), h->devE1.d(), h->devE2.d(), h->vn.d(), h->stream));
    h->refresh_vnw();
    h->velocityParts = 7;
    h->hasVelocity = true;
  });
}

int festung_set_source_points_dev(festung_problem* h, const double* fq, const double* cDe, const double* gNe) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (!fq && !cDe && !gNe) {
      h->set_has_rhs(false);
      return;
    }
    if (cDe) h->require_velocity();
    const int K = h->K, RE = h->tabs.RE, Re3 = 3 * h->tabs.Re;
    if (!h->rhs.p) {
      sync(h);
      h->rhs.alloc(size_t(h->N) * K * sizeof(double));
    }
    h->set_has_rhs(true);
    if (fq) {
      if (!h->devQ2.p) h->devQ2.alloc(size_t(RE) * K * sizeof(double));
      HIPCHK(launch_permute(K, RE, fq, h->dEperm.i(), h->devQ2.d(), h->stream));
    }
    if (cDe) {
      if (!h->devE1.p) h->devE1.alloc(size_t(Re3) * K * sizeof(double));
      HIPCHK(launch_permute(K, Re3, cDe, h->dEperm.i(), h->devE1.d(), h->stream));
    }
    if (gNe) {
      if (!h->devE2.p) h->devE2.alloc(size_t(Re3) * K * sizeof(double));
      HIPCHK(launch_permute(K, Re3, gNe, h->dEperm.i(), h->devE2.d(), h->stream));
    }
    HIPCHK(launch_source(h->p, h->mesh(), fq ? h->devQ2.d() : nullptr, cDe ? h->devE1.d() : nullptr,
                         gNe ? h->devE2.d() : nullptr, h->vn.d(), h->tab.d(), h->off, h->rhs.d(), h->stream));
  });
}

int festung_set_vertex_dirichlet_dev(festung_problem* h, const double* vertexD) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (!(h->flags & FESTUNG_FLAG_LIMITER)) fail(FESTUNG_ERR_STATE, "problem created without the limiter");
    h->set_has_vertexD(vertexD != nullptr);
    if (!vertexD) return;
    HIPCHK(launch_permute(h->numV, 1, vertexD, h->dVperm.i(), h->vertexD.d(), h->stream));
  });
}

int festung_set_source_points(festung_problem* h, const double* fq, const double* cDe, const double* gNe) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (!fq && !cDe && !gNe) {
      h->set_has_rhs(false);
      return;
    }
    if (cDe) h->require_velocity();
    const int K = h->K;
    if (!h->rhs.p) {
      sync(h);
      h->rhs.alloc(size_t(h->N) * K * sizeof(double));
    }
    h->set_has_rhs(true);
    DBuf f, c, gn;
    if (fq) h->upload_elem(f, fq, h->tabs.RE);
    if (cDe) h->upload_elem(c, cDe, 3 * h->tabs.Re);
    if (gNe) h->upload_elem(gn, gNe, 3 * h->tabs.Re);
    HIPCHK(launch_source(h->p, h->mesh(), f.d(), c.d(), gn.d(), h->vn.d(), h->tab.d(), h->off, h->rhs.d(),
                         h->stream));
    sync(h);
  });
}

int festung_set_vertex_dirichlet(festung_problem* h, const double* vertexD) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (!(h->flags & FESTUNG_FLAG_LIMITER)) fail(FESTUNG_ERR_STATE, "problem created without the limiter");
    if (!vertexD) {
      h->set_has_vertexD(false);
      return;
    }
    std::vector<double> v(h->numV);
    for (int32_t d = 0; d < h->numV; ++d) v[d] = vertexD[h->vperm[d]];
    sync(h);  // kernels in flight may still read the old values
    HIPCHK(hipMemcpy(h->vertexD.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    h->set_has_vertexD(true);
  });
}

int festung_project(festung_problem* h, int ord, const double* fq, double* cDisc_out, int set_state) {
  return guard([&] {
    check_handle(h);
    if (!fq) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    bool ordF;
    if (ord == h->tabs.ordE) ordF = false;
    else if (ord == h->tabs.ordF) ordF = true;
    else fail(FESTUNG_ERR_ARG, "projection order %d not available (use %d or %d)", ord, h->tabs.ordE, h->tabs.ordF);
    const int R = ordF ? h->tabs.RF : h->tabs.RE;
    DBuf q, o;
    h->upload_elem(q, fq, R);
    o.alloc(size_t(h->N) * h->K * sizeof(double));
    HIPCHK(launch_project(h->p, ordF, h->K, q.d(), h->tab.d(), h->off, o.d(), h->stream));
    sync(h);
    if (cDisc_out) h->download_elem(o, o.d(), h->N, cDisc_out);
    if (set_state) {
      h->finish_stage_sequence();
      HIPCHK(hipMemcpy(h->X.p, o.p, o.bytes, hipMemcpyDeviceToDevice));
      h->cur = 0;
    }
  });
}

int festung_set_state(festung_problem* h, const double* c) {
  return guard([&] {
    check_handle(h);
    if (!c) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    h->finish_stage_sequence();
    h->upload_elem(h->X, c, h->N, false);
    h->cur = 0;
  });
}

int festung_get_state(festung_problem* h, double* c) {
  return guard([&] {
    check_handle(h);
    if (!c) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    sync(h);
    h->download_elem(h->X, h->buf(h->cur), h->N, c);
  });
}

int festung_set_state_dev(festung_problem* h, const double* d) {
  return guard([&] {
    check_handle(h);
    if (!d) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    h->finish_stage_sequence();
    HIPCHK(launch_permute(h->K, h->N, d, h->dEperm.i(), h->X.d(), h->stream));
    h->cur = 0;
  });
}

int festung_get_state_dev(festung_problem* h, double* d) {
  return guard([&] {
    check_handle(h);
    if (!d) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    HIPCHK(launch_permute(h->K, h->N, h->buf(h->cur), h->dEinv.i(), d, h->stream));
  });
}

int festung_begin_step(festung_problem* h) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (h->nextStage >= 0) fail(FESTUNG_ERR_STATE, "previous step not finished");
    if (h->cur != 0) fail(FESTUNG_ERR_STATE, "state not in the step buffer");
    h->nextStage = 0;
  });
}

int festung_stage(festung_problem* h, int nSubStep, double tau) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (h->nextStage < 0) fail(FESTUNG_ERR_STATE, "festung_begin_step not called");
    h->require_velocity();
    if (nSubStep != h->nextStage + 1) fail(FESTUNG_ERR_ARG, "expected sub-step %d, got %d", h->nextStage + 1, nSubStep);
    const StagePlan& sp = h->plan[nSubStep - 1];
    h->enqueue_stage(sp, tau);
    h->cur = sp.out;
    if (nSubStep == int(h->plan.size())) {
      if (sp.out != 0)
        HIPCHK(hipMemcpyAsync(h->X.p, h->buf(sp.out), size_t(h->N) * h->K * sizeof(double), hipMemcpyDeviceToDevice,
                              h->stream));
      h->cur = 0;
      h->nextStage = -1;
    } else {
      h->nextStage = nSubStep;
    }
  });
}

int festung_step(festung_problem* h, int nsteps, double tau) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    h->require_velocity();
    h->finish_stage_sequence();
    if (nsteps < 0) fail(FESTUNG_ERR_ARG, "negative step count");
    if (nsteps == 0) return;
    if ((h->flags & FESTUNG_FLAG_GRAPH) && !h->prof) {
      if (!(h->graphTau == tau) && (h->graphExec[0] || h->graphExec[1])) h->drop_graph_synced();
      // nsteps = q G + r: the G-step sequence q times, the one-step one r times
      constexpr int G = festung_problem::kGraphSteps;
      const int q = nsteps / G, r = nsteps % G;
      auto captured = [&](int i) -> hipGraphExec_t {
        if (!h->graphExec[i]) {
          HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
          try {
            for (int j = 0; j < (i ? G : 1); ++j) h->enqueue_step(tau);
          } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(h->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          HIPCHK(hipStreamEndCapture(h->stream, &h->graph[i]));
          HIPCHK(hipGraphInstantiate(&h->graphExec[i], h->graph[i], nullptr, nullptr, 0));
          h->graphTau = tau;
        }
        return h->graphExec[i];
      };
      if (q > 0) {
        hipGraphExec_t e = captured(1);
        for (int s = 0; s < q; ++s) HIPCHK(hipGraphLaunch(e, h->stream));
      }
      if (r > 0) {
        hipGraphExec_t e = captured(0);
        for (int s = 0; s < r; ++s) HIPCHK(hipGraphLaunch(e, h->stream));
      }
    } else {
      for (int s = 0; s < nsteps; ++s) h->enqueue_step(tau);
    }
    h->cur = 0;
  });
}

int festung_limit(festung_problem* h) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (!(h->flags & FESTUNG_FLAG_LIMITER)) fail(FESTUNG_ERR_STATE, "problem created without the limiter");
    h->finish_stage_sequence();
    // 'linear' (p >= 1): on the DG coefficients in place, as in the step
    // (limit_dg_kernel): an element the limiter keeps is left as it is, and a
    // limited one is rebuilt from its Taylor mean and gradient.  The change
    // of basis and back that the Taylor path applies to every element scales
    // its terms of degree d with up to cond(B)^d, so on badly shaped elements
    // at p = 4 it returned the coefficients with errors of 1e-10 (a sliver
    // with cond(B) = 188), where the reference's mass-matrix solves keep them.
    const bool dg = h->limType == LT_LINEAR && h->p >= 1;
    if (dg) {
      HIPCHK(launch_centroid(h->K, h->bs, h->X.d(), h->tabs.F[0], h->T2.d(), MinMaxSink{h->partials.d(), 1},
                             h->stream));
      h->enqueue_vertex(h->T2.d(), 1);
    } else {
      ToTaylorArgs ta{h->K, h->B.d(), h->X.d(), h->tab.d(), h->off, h->T2.d(), h->sink(), h->bs};
      HIPCHK(launch_to_taylor(h->p, ta, h->stream));
      h->enqueue_vertex(h->T2.d());
    }
    LimitArgs l{};
    l.m = h->mesh();
    l.taylor = h->T2.d();
    l.vmin = h->vmin.d();
    l.vmax = h->vmax.d();
    l.gmm = h->gmm.d();
    l.vertexD = h->hasVertexD ? h->vertexD.d() : nullptr;
    l.tab = h->tab.d();
    l.off = h->off;
    l.out = h->X.d();
    l.sink = h->sink();
    l.alphaOut = h->alpha.d();
    if (dg) l.taylor = h->X.d();
    HIPCHK(launch_limit(h->p, dg ? LIM_SOLUTION_DG : LIM_SOLUTION, h->limType, l, h->stream));
    sync(h);
    h->cur = 0;
  });
}

int festung_limit_disc_linear(festung_problem* h) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (!(h->flags & FESTUNG_FLAG_LIMITER)) fail(FESTUNG_ERR_STATE, "problem created without the limiter");
    h->finish_stage_sequence();
    // centroid values c_1 phi_1(1/3, 1/3) -> T2 (field 0) and their partials
    HIPCHK(launch_centroid(h->K, h->bs, h->X.d(), phi(1, 1.0 / 3.0, 1.0 / 3.0), h->T2.d(), MinMaxSink{h->partials.d(), 1},
                           h->stream));
    h->enqueue_vertex(h->T2.d(), 1);
    DiscLinArgs l{};
    l.m = h->mesh();
    l.cen = h->T2.d();
    l.vmin = h->vmin.d();
    l.vmax = h->vmax.d();
    l.gmm = h->gmm.d();
    l.vertexD = h->hasVertexD ? h->vertexD.d() : nullptr;
    // the reference element's vertices (0,0), (1,0), (0,1),
    // core/projectDataDisc2DataLagr.m:50
    const double L1[3] = {0.0, 1.0, 0.0}, L2[3] = {0.0, 0.0, 1.0};
    for (int n = 0; n < 3; ++n)
      for (int j = 0; j < 3; ++j) l.phiV[n][j] = phi(j + 1, L1[n], L2[n]);
    l.c = h->X.d();
    l.alphaOut = h->alpha.d();
    HIPCHK(launch_limit_disc_linear(h->p, l, h->stream));
    sync(h);
    h->cur = 0;
  });
}

int festung_rhs(festung_problem* h, double* cDot) {
  return guard([&] {
    check_handle(h);
    if (!cDot) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    h->require_velocity();
    DBuf o;
    o.alloc(size_t(h->N) * h->K * sizeof(double));
    StageArgs a{};
    a.m = h->mesh();
    a.c = h->buf(h->cur);
    a.c0 = h->X.d();
    a.u1 = h->u1.d();
    a.u2 = h->u2.d();
    a.uq = h->uq.d();
    a.vnw = h->vnw.d();
    a.rhs = h->hasRhs ? h->rhs.d() : nullptr;
    a.tab = h->tab.d();
    a.off = h->off;
    a.out = o.d();
    a.sink = h->sink();
    HIPCHK(launch_stage(h->p, OUT_DOT_DG, a, h->stream));
    sync(h);
    h->download_elem(o, o.d(), h->N, cDot);
  });
}

int festung_limiter_info(festung_problem* h, double* alphaE, double* minV0T, double* maxV0T) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if (!(h->flags & FESTUNG_FLAG_LIMITER)) fail(FESTUNG_ERR_STATE, "problem created without the limiter");
    sync(h);
    const int K = h->K;
    if (alphaE) h->download_elem(h->alpha, h->alpha.d(), 1, alphaE);
    if (minV0T || maxV0T) {
      // bounds of the last (solution) limiter application, as the limit
      // kernel forms them: shifted patch min / max joined with c_D
      std::vector<int32_t> v0t(3 * size_t(K));   // device vertex ids, reference element order
      for (size_t i = 0; i < v0t.size(); ++i) v0t[i] = h->vinv[h->V0Tref[i]];
      std::vector<double> vmn(h->numV), vmx(h->numV), vD(h->numV, std::nan("")), gm(2);
      HIPCHK(hipMemcpy(vmn.data(), h->vmin.p, vmn.size() * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(vmx.data(), h->vmax.p, vmx.size() * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(gm.data(), h->gmm.p, 2 * sizeof(double), hipMemcpyDeviceToHost));
      if (h->hasVertexD)
        HIPCHK(hipMemcpy(vD.data(), h->vertexD.p, vD.size() * sizeof(double), hipMemcpyDeviceToHost));
      volatile double sNeg = std::fabs(gm[1]) + 1.0, sPos = std::fabs(gm[0]) + 1.0;
      for (size_t i = 0; i < v0t.size(); ++i) {
        volatile double lo = vmn[v0t[i]] - sNeg;
        lo = lo + sNeg;
        volatile double hi = vmx[v0t[i]] + sPos;
        hi = hi - sPos;
        if (minV0T) minV0T[i] = std::fmin(lo, vD[v0t[i]]);
        if (maxV0T) maxV0T[i] = std::fmax(hi, vD[v0t[i]]);
      }
    }
  });
}

int festung_l2_error(festung_problem* h, int ord, const double* cq, double* err) {
  return guard([&] {
    check_handle(h);
    if (!cq || !err) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    bool ordF;
    if (ord == h->tabs.ordE) ordF = false;
    else if (ord == h->tabs.ordF) ordF = true;
    else fail(FESTUNG_ERR_ARG, "L2 error order %d not available (use %d or %d)", ord, h->tabs.ordE, h->tabs.ordF);
    const int R = ordF ? h->tabs.RF : h->tabs.RE;
    DBuf q, part;
    h->upload_elem(q, cq, R);
    part.alloc(size_t(h->nred) * sizeof(double));
    HIPCHK(launch_l2(h->p, ordF, h->K, h->B.d(), h->buf(h->cur), q.d(), h->tab.d(), h->off, part.d(), h->stream));
    sync(h);
    std::vector<double> hp(h->nred);
    HIPCHK(hipMemcpy(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0;
    for (double v : hp) s += v;
    *err = std::sqrt(s);
  });
}

int festung_set_tau_field(festung_problem* h, const double* tauK) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    h->drop_graph_synced();
    ++h->version;
    if (!tauK) {
      h->hasTauK = false;
      return;
    }
    h->upload_elem(h->tauK, tauK, 1, !h->tauK.p);
    h->hasTauK = true;
  });
}

int festung_set_elem_block(festung_problem* h, int32_t bs) {
  return guard([&] {
    check_handle(h);
    if (!valid_elem_block(bs)) fail(FESTUNG_ERR_ARG, "elements per workgroup must be 64, 128 or 256 (got %d)", bs);
    h->set_device();
    h->finish_stage_sequence();
    if (bs == h->bs) return;
    sync(h);
    h->drop_graph();
    ++h->version;
    h->bs = bs;
    h->nblocks = (h->K + bs - 1) / bs;
  });
}

int festung_get_elem_block(const festung_problem* h, int32_t* bs) {
  return guard([&] {
    check_handle(h);
    if (!bs) fail(FESTUNG_ERR_ARG, "null pointer argument");
    *bs = h->bs;
  });
}

int festung_autotune(festung_problem* h, double tau, int32_t nsteps, int32_t* best, double* msPerStep) {
  return guard([&] {
    check_handle(h);
    if (nsteps < 1) fail(FESTUNG_ERR_ARG, "autotune needs at least one timed step");
    h->set_device();
    h->require_velocity();
    h->finish_stage_sequence();
    if (h->prof) fail(FESTUNG_ERR_STATE, "profiling is on");
    // the trial steps advance the state: keep it and put it back afterwards
    DBuf save;
    const size_t nk = size_t(h->N) * h->K * sizeof(double);
    save.alloc(nk);
    HIPCHK(hipMemcpyAsync(save.p, h->buf(h->cur), nk, hipMemcpyDeviceToDevice, h->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const int cand[3] = {64, 128, 256};
    double bestMs = 0.0;
    int bestBs = h->bs;
    try {
      for (int c = 0; c < 3; ++c) {
        check_call(festung_set_elem_block(h, cand[c]));
        check_call(festung_step(h, 2, tau));  // capture + warm-up
        HIPCHK(hipEventRecord(e0, h->stream));
        check_call(festung_step(h, nsteps, tau));
        HIPCHK(hipEventRecord(e1, h->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        const double per = double(ms) / nsteps;
        if (msPerStep) msPerStep[c] = per;
        if (c == 0 || per < bestMs) {
          bestMs = per;
          bestBs = cand[c];
        }
        HIPCHK(hipMemcpyAsync(h->X.p, save.p, nk, hipMemcpyDeviceToDevice, h->stream));
        h->cur = 0;
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_call(festung_set_elem_block(h, bestBs));
    sync(h);
    tune_record(h->tune_key(), bestBs);
    if (best) *best = bestBs;
  });
}

int festung_adj_rhs(festung_problem* h, const double* mu, double* out) {
  return guard([&] {
    check_handle(h);
    if (!mu || !out) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    h->require_velocity();
    DBuf dm, o;
    h->upload_elem(dm, mu, h->N);
    o.alloc(size_t(h->N) * h->K * sizeof(double));
    AdjArgs a{};
    a.m = h->mesh();
    a.mu = dm.d();
    a.u1 = h->u1.d();
    a.u2 = h->u2.d();
    a.vn = h->vn.d();
    a.vnw = h->vnw.d();
    a.tab = h->tab.d();
    a.off = h->off;
    a.mode = 4;
    a.out = o.d();
    HIPCHK(launch_adj_stage(h->p, a, h->stream));
    sync(h);
    h->download_elem(o, o.d(), h->N, out);
  });
}

// Discrete adjoint with binomial checkpointing: only state snapshots are
// kept (`snaps` of them, the initial state included), placed by the binomial
// schedule; each step is recomputed with its stage states right before it
// is reversed.  Velocity sensitivities fused into the transposed stage (or
// split over three threads per element); one stream.  The stage kernels and
// their order are those of the uniform schedule, so the results are the same
// bits as festung_adjoint's with fused (split) sensitivities.
static void adjoint_binomial(festung_problem* h, int nsteps, double tau, const double* cObs, int snaps, double* J,
                             double* gradC0, double* gradU1, double* gradU2, double* gradVn, double* timing_ms) {
  if (snaps < 1) fail(FESTUNG_ERR_ARG, "binomial checkpointing needs at least one snapshot");
  if (h->flags & FESTUNG_FLAG_ADJ_DEFER)
    fail(FESTUNG_ERR_ARG, "binomial checkpointing forms the sensitivities fused or split, not deferred");
  std::vector<CkAction> plan;
  try {
    plan = binomial_schedule(nsteps, snaps);
  } catch (const std::exception& e) {
    fail(FESTUNG_ERR_INTERNAL, "%s", e.what());
  }
  const int K = h->K, N = h->N, Re = h->tabs.Re, S = int(h->plan.size());
  const size_t nk = size_t(N) * K * sizeof(double);
  const int what = (h->flags & FESTUNG_FLAG_ADJ_SPLIT) ? ADJ_SPLIT : ADJ_FUSED;
  hipStream_t A = h->stream;
  const int c = std::min(snaps, nsteps);
  std::vector<std::unique_ptr<DBuf>> snap(c);
  for (auto& b : snap) {
    b.reset(new DBuf);
    b->alloc(nk);
  }
  // rot: advance ping-pong; rotOut: the schedule's scratch state; tmp: an
  // advance's stage outputs; stg: the stage states of the step being
  // reversed; fin: the final state; junk: other recomputed step ends
  DBuf rot[2], rotOut, fin, junk, dObs, part, L0, Gu1, Gu2, Gvn, mass;
  std::vector<DBuf> tmp(std::max(S - 1, 1)), stg(std::max(S - 1, 1));
  for (DBuf* b : {&rot[0], &rot[1], &rotOut, &fin, &junk, &L0, &Gu1, &Gu2, &mass}) b->alloc(nk);
  for (auto& b : tmp) b.alloc(nk);
  for (auto& b : stg) b.alloc(nk);
  Gvn.alloc(size_t(3) * Re * K * sizeof(double));
  part.alloc(size_t(h->nred) * sizeof(double));
  h->load_in(dObs, cObs, N);
  constexpr int R = 4;  // ring of adjoint stage states
  DBuf ringb[R];
  for (auto& b : ringb) b.alloc(nk);
  auto ring = [&](long j) -> double* { return ringb[j % R].d(); };
  std::vector<hipEvent_t> ev(3, nullptr);
  struct EvGuard {
    std::vector<hipEvent_t>& v;
    ~EvGuard() {
      for (auto e : v)
        if (e) (void)hipEventDestroy(e);
    }
  } evg{ev};
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));

  // one forward step in -> out; stage outputs into outs[0 .. S-2]
  auto fwd_step = [&](const double* in, double* out, std::vector<DBuf>& outs) {
    for (int st = 0; st < S; ++st) {
      StageArgs a{};
      a.m = h->mesh();
      a.c = st == 0 ? in : outs[st - 1].d();
      a.c0 = in;
      a.out = st + 1 < S ? outs[st].d() : out;
      a.u1 = h->u1.d();
      a.u2 = h->u2.d();
      a.uq = h->uq.d();
      a.vnw = h->vnw.d();
      a.tab = h->tab.d();
      a.off = h->off;
      a.tau = tau;
      a.tauK = h->hasTauK ? h->tauK.d() : nullptr;
      a.rk = h->plan[st].rk;
      a.sink = h->sink();
      HIPCHK(launch_stage(h->p, OUT_RK_DG, a, A));
    }
  };
  auto buf = [&](int32_t id) -> double* { return id == kRot ? rotOut.d() : snap[id]->d(); };
  HIPCHK(hipEventRecord(ev[0], A));
  HIPCHK(hipMemcpyAsync(snap[0]->p, h->buf(h->cur), nk, hipMemcpyDeviceToDevice, A));
  long i = 0;  // adjoint stage counter
  for (const CkAction& x : plan) {
    if (x.kind == CkAction::ADVANCE) {
      const double* cur = buf(x.src);
      for (int l = 0; l < x.steps; ++l) {
        double* next = l + 1 == x.steps ? buf(x.dst) : rot[l % 2].d();
        fwd_step(cur, next, tmp);
        cur = next;
      }
      continue;
    }
    // REVERSE step j: recompute it with its stage states, then its
    // transposed stages in reverse order
    const int j = x.from;
    const double* in = buf(x.src);
    fwd_step(in, j + 1 == nsteps ? fin.d() : junk.d(), stg);
    if (j + 1 == nsteps) {
      // loss and terminal adjoint (into ring[R-1], read by the first stage)
      HIPCHK(hipEventRecord(ev[1], A));
      HIPCHK(launch_loss(h->p, K, h->B.d(), fin.d(), dObs.d(), h->tab.d(), h->off, part.d(), ring(R - 1), A));
      HIPCHK(launch_fill(size_t(N) * K, 0.0, Gu1.d(), A));
      HIPCHK(launch_fill(size_t(N) * K, 0.0, Gu2.d(), A));
      HIPCHK(launch_fill(size_t(3) * Re * K, 0.0, Gvn.d(), A));
    }
    for (int st = S - 1; st >= 0; --st, ++i) {
      AdjArgs a{};
      a.m = h->mesh();
      a.mu = ring(i + R - 1);
      a.u1 = h->u1.d();
      a.u2 = h->u2.d();
      a.vn = h->vn.d();
      a.vnw = h->vnw.d();
      a.tab = h->tab.d();
      a.off = h->off;
      a.tau = tau;
      a.tauK = h->hasTauK ? h->tauK.d() : nullptr;
      a.omega = h->plan[st].rk.a0;
      a.L0 = L0.d();
      a.mode = S == 1 ? 3 : (st == S - 1 ? 0 : (st > 0 ? 1 : 2));
      a.out = ring(i);
      a.c = st == 0 ? in : stg[st - 1].d();
      a.Gu1 = Gu1.d();
      a.Gu2 = Gu2.d();
      a.Gvn = Gvn.d();
      a.what = what;
      HIPCHK(launch_adj_stage(h->p, a, A));
    }
  }
  HIPCHK(launch_mass(h->p, K, h->B.d(), ring(i + R - 1), h->tab.d(), h->off, mass.d(), A));
  HIPCHK(hipEventRecord(ev[2], A));
  HIPCHK(hipMemcpyAsync(h->X.p, fin.p, nk, hipMemcpyDeviceToDevice, A));
  h->cur = 0;
  sync(h);
  std::vector<double> hp(h->nred);
  HIPCHK(hipMemcpy(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost));
  double jsum = 0;
  for (double v : hp) jsum += v;
  if (J) *J = jsum;
  if (gradC0) h->store_out(mass, N, gradC0);
  if (gradU1) h->store_out(Gu1, N, gradU1);
  if (gradU2) h->store_out(Gu2, N, gradU2);
  if (gradVn) h->store_out(Gvn, 3 * Re, gradVn);
  if (timing_ms) {
    float f = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&f, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
    timing_ms[0] = f;
    timing_ms[1] = b;
  }
}

int festung_adjoint(festung_problem* h, int32_t nsteps, double tau, const double* cObs, int32_t ckpt, double* J,
                    double* gradC0, double* gradU1, double* gradU2, double* gradVn, double* timing_ms) {
  return guard([&] {
    check_handle(h);
    if (!cObs || nsteps < 1) fail(FESTUNG_ERR_ARG, "invalid argument");
    h->set_device();
    h->require_velocity();
    h->finish_stage_sequence();
    if (h->flags & FESTUNG_FLAG_LIMITER)
      fail(FESTUNG_ERR_UNSUPPORTED, "the adjoint needs the limiter off (the limiter is not differentiable)");
    if (h->ordRK > 3) fail(FESTUNG_ERR_UNSUPPORTED, "the adjoint supports the SSP orders 1..3");
    if (h->hasRhs) fail(FESTUNG_ERR_UNSUPPORTED, "the adjoint assumes f = 0 and c_D = 0");
    const int K = h->K, N = h->N, Re = h->tabs.Re, S = int(h->plan.size());
    const size_t nk = size_t(N) * K * sizeof(double);
    if (h->flags & FESTUNG_FLAG_ADJ_BINOMIAL) {
      adjoint_binomial(h, nsteps, tau, cObs, ckpt, J, gradC0, gradU1, gradU2, gradVn, timing_ms);
      return;
    }
    const int C = (ckpt <= 0 || ckpt > nsteps) ? nsteps : ckpt;
    const int nseg = (nsteps + C - 1) / C;
    // Velocity sensitivities: fused into the transposed stage kernel (one
    // thread per element), split over three threads per element, or deferred
    // to one pass per checkpoint segment after its sweep (the sweep's kernels
    // then only apply the transposed stages; the segment's adjoint stage
    // inputs are kept for the pass).  Small meshes cannot fill the GPU with
    // one thread per element: there the sweep's critical path is what counts.
    int what = h->nblocks < 2048 ? ADJ_DEFER : ADJ_FUSED;
    if (h->flags & FESTUNG_FLAG_ADJ_FUSED) what = ADJ_FUSED;
    if (h->flags & FESTUNG_FLAG_ADJ_SPLIT) what = ADJ_SPLIT;
    if (h->flags & FESTUNG_FLAG_ADJ_DEFER) what = ADJ_DEFER;
    const bool defer = what == ADJ_DEFER;
    // Streams: A = the problem's stream (the reverse sweep); Cs = the
    // recomputation of the next segment's stage states into another buffer
    // set, overlapping the sweep (unless FESTUNG_FLAG_ADJ_SERIAL); Ss = the
    // deferred sensitivity passes, overlapping both.  A buffer set is reused
    // once its segment is done (swept, and its sensitivity pass run): two
    // sets, three with deferred sensitivities.
    const bool overlap = !(h->flags & FESTUNG_FLAG_ADJ_SERIAL) && nseg > 2;
    hipStream_t A = h->stream, Cst = A, Sst = A;
    if (overlap) {
      h->ensure_aux_streams();
      Cst = h->aux[0];
    }
    // (with a single segment -- all states stored -- the passes overlap its sweep)
    if (defer && !(h->flags & FESTUNG_FLAG_ADJ_SERIAL) && (overlap || nseg == 1)) {
      h->ensure_aux_streams();
      Sst = h->aux[1];
    }
    const int nsets = overlap ? (defer ? 3 : 2) : 1;
    std::vector<std::unique_ptr<DBuf>> ck(nseg + 1), sb(size_t(nsets) * C * S);
    for (auto& b : ck) {
      b.reset(new DBuf);
      b->alloc(nk);
    }
    for (size_t j = 0; j < sb.size(); ++j)
      if (j % (size_t(C) * S) != 0) {  // stage 0 of local step 0 lives in the checkpoint
        sb[j].reset(new DBuf);
        sb[j]->alloc(nk);
      }
    // ring of adjoint stage states: stage i reads ring[(i-1) mod R], writes
    // ring[i mod R]; deferred sensitivities keep the inputs of as many
    // segments as may be in flight (a segment's pass can overlap the sweeps
    // of the next nsets - 1 segments)
    const int R = defer ? nsets * C * S + 2 : 4;
    std::vector<std::unique_ptr<DBuf>> ringb(R);
    for (auto& b : ringb) {
      b.reset(new DBuf);
      b->alloc(nk);
    }
    auto ring = [&](long j) -> double* { return ringb[size_t(j % R)]->d(); };
    DBuf dObs, part, L0, Gu1, Gu2, Gvn, tmp;
    h->load_in(dObs, cObs, N);
    part.alloc(size_t(h->nred) * sizeof(double));
    for (DBuf* b : {&L0, &Gu1, &Gu2, &tmp}) b->alloc(nk);
    Gvn.alloc(size_t(3) * Re * K * sizeof(double));
    // events: 0..2 timing; segment hand-offs between A and Cs
    std::vector<hipEvent_t> ev(3, nullptr), evSeg(nsets, nullptr), evDone(nsets, nullptr), evFwd(1, nullptr),
        evSwept(nsets, nullptr), evFin(1, nullptr);
    struct EvGuard {
      std::vector<hipEvent_t>* v[6];
      ~EvGuard() {
        for (auto* x : v)
          for (auto e : *x)
            if (e) (void)hipEventDestroy(e);
      }
    } evg{{&ev, &evSeg, &evDone, &evFwd, &evSwept, &evFin}};
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    for (auto* x : {&evSeg, &evDone, &evFwd, &evSwept, &evFin})
      for (auto& e : *x) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    auto state = [&](int seg, int l, int len, int st) -> double* {
      if (l == len) return ck[seg + 1]->d();
      if (l == 0 && st == 0) return ck[seg]->d();
      return sb[(size_t(seg % nsets) * C + l) * S + st]->d();
    };
    // A recomputation does not rewrite the segment's end checkpoint (the
    // sweep may read it meanwhile): its last stage output goes to junk.
    // Segments whose stage states are not kept run through a small rotation
    // of buffers instead (only their end checkpoint is written).
    DBuf junk, rot[4];
    if (nseg > 1) junk.alloc(nk);
    if (nseg > nsets)
      for (auto& b : rot) b.alloc(nk);
    enum { SEG_STORE, SEG_RECOMPUTE, SEG_PASS };
    auto run_segment = [&](int seg, hipStream_t sm, int how) {
      const int len = std::min(C, nsteps - seg * C);
      double* cur = ck[seg]->d();
      for (int l = 0; l < len; ++l) {
        double* next = l + 1 == len ? ck[seg + 1]->d() : rot[l % 2].d();
        for (int st = 0; st < S; ++st) {
          StageArgs a{};
          a.m = h->mesh();
          if (how == SEG_PASS) {
            a.c = st == 0 ? cur : rot[1 + st].d();
            a.c0 = cur;
            a.out = st + 1 < S ? rot[2 + st].d() : next;
          } else {
            a.c = state(seg, l, len, st);
            a.c0 = state(seg, l, len, 0);
            a.out = st + 1 < S ? state(seg, l, len, st + 1) : state(seg, l + 1, len, 0);
            if (how == SEG_RECOMPUTE && l + 1 == len && st + 1 == S) a.out = junk.d();
          }
          a.u1 = h->u1.d();
          a.u2 = h->u2.d();
          a.uq = h->uq.d();
          a.vnw = h->vnw.d();
          a.tab = h->tab.d();
          a.off = h->off;
          a.tau = tau;
          a.tauK = h->hasTauK ? h->tauK.d() : nullptr;
          a.rk = h->plan[st].rk;
          a.sink = h->sink();
          HIPCHK(launch_stage(h->p, OUT_RK_DG, a, sm));
        }
        cur = next;
      }
    };

    // deferred sensitivities: per adjoint stage (sweep order) its input mu,
    // forward stage state and omega, on the device for the segment passes.
    // A pass sums the stages of each time step as one chunk and adds the
    // chunk sums in sweep order: the same order for any checkpoint interval.
    // At most nChunk steps per pass (the sweep's end waits for the last one).
    const int nChunk = 32;
    DBuf dMus, dCs, dOm;
    if (defer) {
      std::vector<const double*> hm, hc;
      std::vector<double> ho;
      long j = 0;
      for (int seg = nseg - 1; seg >= 0; --seg) {
        const int len = std::min(C, nsteps - seg * C);
        for (int l = len - 1; l >= 0; --l)
          for (int st = S - 1; st >= 0; --st, ++j) {
            hm.push_back(ring(j + R - 1));
            hc.push_back(state(seg, l, len, st));
            ho.push_back(h->plan[st].rk.a0);
          }
      }
      h->upload(dMus, hm.data(), hm.size());
      h->upload(dCs, hc.data(), hc.size());
      h->upload(dOm, ho.data(), ho.size());
    }
    // forward, storing checkpoints (and the stage states of the last segments)
    HIPCHK(hipEventRecord(ev[0], A));
    HIPCHK(hipMemcpyAsync(ck[0]->p, h->buf(h->cur), nk, hipMemcpyDeviceToDevice, A));
    for (int seg = 0; seg < nseg; ++seg) run_segment(seg, A, seg < nseg - nsets ? SEG_PASS : SEG_STORE);
    HIPCHK(hipEventRecord(ev[1], A));
    // loss and terminal adjoint (into ring[R-1], read by stage 0)
    HIPCHK(launch_loss(h->p, K, h->B.d(), ck[nseg]->d(), dObs.d(), h->tab.d(), h->off, part.d(), ring(R - 1), A));
    HIPCHK(launch_fill(size_t(N) * K, 0.0, Gu1.d(), A));
    HIPCHK(launch_fill(size_t(N) * K, 0.0, Gu2.d(), A));
    HIPCHK(launch_fill(size_t(3) * Re * K, 0.0, Gvn.d(), A));
    if (overlap) {
      HIPCHK(hipEventRecord(evFwd[0], A));
      // enqueue the second stream's work only once the forward pass is done:
      // a pending cross-stream wait slows the forward's dispatch
      HIPCHK(hipEventSynchronize(ev[1]));
      HIPCHK(hipStreamWaitEvent(Cst, evFwd[0], 0));
    }
    long i = 0;  // adjoint stage counter
    for (int seg = nseg - 1; seg >= 0; --seg) {
      const long segFirst = i;
      const int len = std::min(C, nsteps - seg * C);
      // The forward pass leaves the stage states of the last nsets segments
      // in place; earlier ones are recomputed from their checkpoints.
      if (seg < nseg - nsets) {
        if (overlap) HIPCHK(hipStreamWaitEvent(A, evSeg[seg % nsets], 0));
        else run_segment(seg, A, SEG_RECOMPUTE);
      }
      // recompute the previous segment on Cs, into its set, once the segment
      // that used that set last (seg - 1 + nsets) is done
      if (overlap && seg >= 1 && seg - 1 < nseg - nsets) {
        HIPCHK(hipStreamWaitEvent(Cst, evDone[(seg - 1) % nsets], 0));
        run_segment(seg - 1, Cst, SEG_RECOMPUTE);
        HIPCHK(hipEventRecord(evSeg[(seg - 1) % nsets], Cst));
      }
      // deferred sensitivities: a pass over every nChunk swept steps (in
      // sweep order), on Ss once those steps are swept
      auto sens_pass = [&](long first, int nsteps_) {
        if (Sst != A) {
          HIPCHK(hipEventRecord(evSwept[seg % nsets], A));
          HIPCHK(hipStreamWaitEvent(Sst, evSwept[seg % nsets], 0));
        }
        SensArgs sa{};
        sa.m = h->mesh();
        sa.vn = h->vn.d();
        sa.tab = h->tab.d();
        sa.off = h->off;
        sa.tau = tau;
        sa.tauK = h->hasTauK ? h->tauK.d() : nullptr;
        sa.mus = static_cast<const double* const*>(dMus.p);
        sa.cs = static_cast<const double* const*>(dCs.p);
        sa.omegas = dOm.d();
        sa.Gu1 = Gu1.d();
        sa.Gu2 = Gu2.d();
        sa.Gvn = Gvn.d();
        sa.first = int(first);
        sa.count = nsteps_ * S;
        sa.nchunk = nsteps_;
        HIPCHK(launch_sens_segment(h->p, sa, Sst));
      };
      long passFirst = segFirst;
      int passSteps = 0;
      for (int l = len - 1; l >= 0; --l) {
        for (int st = S - 1; st >= 0; --st, ++i) {
          AdjArgs a{};
          a.m = h->mesh();
          a.mu = ring(i + R - 1);
          a.u1 = h->u1.d();
          a.u2 = h->u2.d();
          a.vn = h->vn.d();
          a.vnw = h->vnw.d();
          a.tab = h->tab.d();
          a.off = h->off;
          a.tau = tau;
          a.tauK = h->hasTauK ? h->tauK.d() : nullptr;
          a.omega = h->plan[st].rk.a0;
          a.L0 = L0.d();
          a.mode = S == 1 ? 3 : (st == S - 1 ? 0 : (st > 0 ? 1 : 2));
          a.out = ring(i);
          a.c = state(seg, l, len, st);
          a.Gu1 = Gu1.d();
          a.Gu2 = Gu2.d();
          a.Gvn = Gvn.d();
          a.what = defer ? ADJ_STAGE : what;
          HIPCHK(launch_adj_stage(h->p, a, A));
        }
        if (defer && (++passSteps == nChunk || l == 0)) {
          sens_pass(passFirst, passSteps);
          passFirst = i;
          passSteps = 0;
        }
      }
      if (overlap) HIPCHK(hipEventRecord(evDone[seg % nsets], defer ? Sst : A));
    }
    if (Sst != A) {  // the last sensitivity pass, before the sweep's end
      HIPCHK(hipEventRecord(evFin[0], Sst));
      HIPCHK(hipStreamWaitEvent(A, evFin[0], 0));
    }
    const double* muFinal = ring(i + R - 1);
    HIPCHK(launch_mass(h->p, K, h->B.d(), muFinal, h->tab.d(), h->off, tmp.d(), A));
    HIPCHK(hipEventRecord(ev[2], A));
    // final state becomes the problem's state
    HIPCHK(hipMemcpyAsync(h->X.p, ck[nseg]->p, nk, hipMemcpyDeviceToDevice, A));
    h->cur = 0;
    sync(h);
    std::vector<double> hp(h->nred);
    HIPCHK(hipMemcpy(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost));
    double j = 0;
    for (double v : hp) j += v;
    if (J) *J = j;
    if (gradC0) h->store_out(tmp, N, gradC0);
    if (gradU1) h->store_out(Gu1, N, gradU1);
    if (gradU2) h->store_out(Gu2, N, gradU2);
    if (gradVn) h->store_out(Gvn, 3 * Re, gradVn);
    if (timing_ms) {
      float f = 0, b = 0;
      HIPCHK(hipEventElapsedTime(&f, ev[0], ev[1]));
      HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
      timing_ms[0] = f;
      timing_ms[1] = b;
    }
  });
}

int festung_adjoint_dev(festung_problem* h, int32_t nsteps, double tau, const double* cObs, int32_t ckpt,
                        double* J, double* gradC0, double* gradU1, double* gradU2, double* gradVn,
                        double* timing_ms) {
  return guard([&] {
    check_handle(h);
    struct Reset {
      festung_problem* h;
      ~Reset() { h->ioDev = false; }
    } reset{h};
    h->ioDev = true;
    check_call(festung_adjoint(h, nsteps, tau, cObs, ckpt, J, gradC0, gradU1, gradU2, gradVn, timing_ms));
  });
}

// One operator of the reference's assembled set applied alone (OpKind).
static void apply_operator(festung_problem* h, int op, const double* x, double* y) {
  if (op < FESTUNG_OP_G || op > FESTUNG_OP_MINV) fail(FESTUNG_ERR_ARG, "unknown operator %d", op);
  if (!x || !y) fail(FESTUNG_ERR_ARG, "null pointer argument");
  h->set_device();
  if (op <= FESTUNG_OP_R) h->require_velocity();
  DBuf dx, dy;
  h->load_in(dx, x, h->N);
  dy.alloc(size_t(h->N) * h->K * sizeof(double));
  HIPCHK(launch_apply(h->p, op, h->mesh(), dx.d(), h->u1.d(), h->u2.d(), h->vnw.d(), h->tab.d(), h->off, dy.d(),
                      h->stream));
  h->store_out(dy, h->N, y);
}

int festung_apply_operator(festung_problem* h, int op, const double* x, double* y) {
  return guard([&] {
    check_handle(h);
    apply_operator(h, op, x, y);
  });
}

int festung_apply_operator_dev(festung_problem* h, int op, const double* x, double* y) {
  return guard([&] {
    check_handle(h);
    struct Reset {
      festung_problem* h;
      ~Reset() { h->ioDev = false; }
    } reset{h};
    h->ioDev = true;
    apply_operator(h, op, x, y);
  });
}

// Scratch of festung_apply_operator_dev_noalloc: the operand and the result
// in the device element order, kNoAllocValues doubles each (64 MiB in all).
// Every device the library runs on holds its own copy until the process
// ends; its address there is looked up once per device (scratch_on).
constexpr size_t kNoAllocValues = size_t(1) << 22;
static __device__ __align__(256) double scratch[2 * kNoAllocValues];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
double* g_scratchPtr[kMaxDevices] = {};

// the scratch's device address on `device` (current): hipGetSymbolAddress the
// first time, the kept pointer afterwards (the host-side name of the array
// is not a device address)
double* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) fail(FESTUNG_ERR_UNSUPPORTED, "device %d: too many devices", device);
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchPtr[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratchPtr[device] = static_cast<double*>(p);
  }
  return g_scratchPtr[device];
}
}  // namespace

int festung_apply_operator_dev_noalloc(festung_problem* h, int op, const double* x, double* y) {
  return guard([&] {
    check_handle(h);
    if (op < FESTUNG_OP_G || op > FESTUNG_OP_MINV) fail(FESTUNG_ERR_ARG, "unknown operator %d", op);
    if (!x || !y) fail(FESTUNG_ERR_ARG, "null pointer argument");
    const size_t kn = size_t(h->K) * h->N;
    if (kn > kNoAllocValues)
      fail(FESTUNG_ERR_ARG, "K*N = %zu exceeds the %zu values of the no-allocation scratch", kn, kNoAllocValues);
    h->set_device();
    if (op <= FESTUNG_OP_R) h->require_velocity();
    double* s = scratch_on(h->device);
    double* dx = s;
    double* dy = s + kNoAllocValues;
    HIPCHK(launch_permute(h->K, h->N, x, h->dEperm.i(), dx, h->stream));
    HIPCHK(launch_apply(h->p, op, h->mesh(), dx, h->u1.d(), h->u2.d(), h->vnw.d(), h->tab.d(), h->off, dy,
                        h->stream));
    HIPCHK(launch_permute(h->K, h->N, dy, h->dEinv.i(), y, h->stream));
  });
}

int festung_assemble_vectors(festung_problem* h, const double* fq, const double* cDe, const double* gNe, double* L,
                             double* KD, double* KN) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    if ((L && !fq) || (KD && !cDe) || (KN && !gNe)) fail(FESTUNG_ERR_ARG, "an output needs its point values");
    if (KD) h->require_velocity();
    const int K = h->K, N = h->N;
    DBuf pts, out;
    out.alloc(size_t(N) * K * sizeof(double));
    // one source launch per vector, from that vector's data alone: globL
    // with the sign it has, globKD and globKN negated back
    auto one = [&](int which, const double* src, int cols, double* dst) {
      h->upload_elem(pts, src, cols);
      HIPCHK(launch_source(h->p, h->mesh(), which == 0 ? pts.d() : nullptr, which == 1 ? pts.d() : nullptr,
                           which == 2 ? pts.d() : nullptr, h->vn.d(), h->tab.d(), h->off, out.d(), h->stream));
      h->download_elem(out, out.d(), N, dst);
      if (which != 0)
        for (size_t i = 0; i < size_t(N) * K; ++i) dst[i] = -dst[i];
    };
    if (L) one(0, fq, h->tabs.RE, L);
    if (KD) one(1, cDe, 3 * h->tabs.Re, KD);
    if (KN) one(2, gNe, 3 * h->tabs.Re, KN);
  });
}

int festung_vertex_based_limiter(festung_problem* h, const double* valCentroid, const double* valV0T,
                                 const uint8_t* markV0TbdrD, const double* dataV0T, double* alphaE, double* minV0T,
                                 double* maxV0T) {
  return guard([&] {
    check_handle(h);
    if (!valCentroid || !valV0T) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    const int K = h->K, numV = h->numV;
    // valD(markV0TbdrD) = dataV0T(markV0TbdrD), NaN elsewhere
    // (core/computeMinMaxV0TElementPatch.m:81-82)
    const bool hasD = markV0TbdrD && dataV0T;
    std::vector<double> valD;
    if (hasD) {
      valD.assign(3 * size_t(K), std::numeric_limits<double>::quiet_NaN());
      for (size_t i = 0; i < valD.size(); ++i)
        if (markV0TbdrD[i]) valD[i] = dataV0T[i];
    }
    DBuf cenIn, cen, val, dD, part, gm, vmn, vmx, al, mn, mx;
    h->upload_elem(cenIn, valCentroid, 1);
    h->upload_elem(val, valV0T, 3);
    if (hasD) h->upload_elem(dD, valD.data(), 3);
    cen.alloc(size_t(K) * sizeof(double));
    part.alloc(2 * size_t(h->nblocks) * sizeof(double));
    gm.alloc(2 * sizeof(double));
    vmn.alloc(size_t(numV) * sizeof(double));
    vmx.alloc(size_t(numV) * sizeof(double));
    al.alloc(size_t(K) * sizeof(double));
    mn.alloc(3 * size_t(K) * sizeof(double));
    mx.alloc(3 * size_t(K) * sizeof(double));
    HIPCHK(launch_centroid(K, h->bs, cenIn.d(), 1.0, cen.d(), MinMaxSink{part.d(), 1}, h->stream));
    VertexArgs va{numV, h->maxDeg, h->v2tPad.i(), cen.d(), K, 1, 1.0, part.d(), h->nblocks, gm.d(), vmn.d(), vmx.d()};
    HIPCHK(launch_vertex(va, h->stream));
    HIPCHK(launch_vertex_limiter(h->mesh(), cen.d(), val.d(), hasD ? dD.d() : nullptr, vmn.d(), vmx.d(), gm.d(),
                                 al.d(), mn.d(), mx.d(), h->stream));
    if (alphaE) h->download_elem(al, al.d(), 1, alphaE);
    if (minV0T) h->download_elem(mn, mn.d(), 3, minV0T);
    if (maxV0T) h->download_elem(mx, mx.d(), 3, maxV0T);
    sync(h);
  });
}

int festung_apply_slope_limiter_disc(festung_problem* h, const double* dataDisc, const uint8_t* markV0TbdrD,
                                     const double* dataV0T, double* dataDiscLim, double* minV0T, double* maxV0T) {
  return guard([&] {
    check_handle(h);
    if (!dataDisc || !dataDiscLim) fail(FESTUNG_ERR_ARG, "null pointer argument");
    if (!(h->flags & FESTUNG_FLAG_LIMITER)) fail(FESTUNG_ERR_STATE, "problem created without the limiter");
    const int K = h->K, numV = h->numV;
    // the per-(element, vertex) Dirichlet data as per-vertex values: they
    // must agree between the elements that share a vertex
    std::vector<double> vD;
    if (markV0TbdrD && dataV0T) {
      vD.assign(numV, std::numeric_limits<double>::quiet_NaN());
      std::vector<int8_t> seen(numV, -1);
      for (size_t i = 0; i < 3 * size_t(K); ++i) {
        const int32_t v = h->V0Tref[i];
        const int8_t m = markV0TbdrD[i] ? 1 : 0;
        if (seen[v] >= 0 && seen[v] != m)
          fail(FESTUNG_ERR_ARG, "markV0TbdrD differs between the elements sharing vertex %d", v);
        seen[v] = m;
        if (!m) continue;
        if (!std::isnan(vD[v]) && !(vD[v] == dataV0T[i]))
          fail(FESTUNG_ERR_ARG, "dataV0T differs between the elements sharing vertex %d", v);
        vD[v] = dataV0T[i];
      }
    }
    check_call(festung_set_state(h, dataDisc));
    check_call(festung_set_vertex_dirichlet(h, vD.empty() ? nullptr : vD.data()));
    check_call(festung_limit(h));
    check_call(festung_get_state(h, dataDiscLim));
    if (minV0T || maxV0T) check_call(festung_limiter_info(h, nullptr, minV0T, maxV0T));
  });
}

int festung_runge_kutta_explicit(int ord, double tau, double t0, double* t, double* omega, int32_t* nStages) {
  return guard([&] {
    if (!t || !omega) fail(FESTUNG_ERR_ARG, "null pointer argument");
    // core/rungeKuttaExplicit.m:64-78 (SSP orders 1..3); 4 = classic RK4,
    // whose stage times are returned with omega = NaN (not a convex form)
    static const double c1[1] = {0.0}, w1[1] = {0.0};
    static const double c2[2] = {0.0, 1.0}, w2[2] = {0.0, 0.5};
    static const double c3[3] = {0.0, 1.0, 0.5}, w3[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
    static const double c4[4] = {0.0, 0.5, 0.5, 1.0};
    const double* c = nullptr;
    const double* w = nullptr;
    int S = 0;
    switch (ord) {
      case 1: c = c1; w = w1; S = 1; break;
      case 2: c = c2; w = w2; S = 2; break;
      case 3: c = c3; w = w3; S = 3; break;
      case 4: c = c4; S = 4; break;
      default: fail(FESTUNG_ERR_UNSUPPORTED, "Runge-Kutta order %d not supported (1..4)", ord);
    }
    for (int i = 0; i < S; ++i) {
      t[i] = t0 + c[i] * tau;
      omega[i] = w ? w[i] : std::numeric_limits<double>::quiet_NaN();
    }
    if (nStages) *nStages = S;
  });
}

int festung_profile(festung_problem* h, int enable) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    sync(h);
    for (auto& r : h->recs) {
      h->pool.push_back(r.a);
      h->pool.push_back(r.b);
    }
    h->recs.clear();
    h->prof = enable != 0;
  });
}

int festung_profile_read(festung_problem* h, double* ms, int64_t* count) {
  return guard([&] {
    check_handle(h);
    if (!ms || !count) fail(FESTUNG_ERR_ARG, "null pointer argument");
    h->set_device();
    sync(h);
    for (int i = 0; i < FESTUNG_KERNEL_KINDS; ++i) {
      ms[i] = 0;
      count[i] = 0;
    }
    for (auto& r : h->recs) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.kind] += t;
      count[r.kind] += 1;
      h->pool.push_back(r.a);
      h->pool.push_back(r.b);
    }
    h->recs.clear();
  });
}

}  // extern "C"

// Several problems advanced together: one captured hipGraph whose branches
// are the problems' step sequences (each on its own stream), launched once.
struct festung_batch {
  std::vector<festung_problem*> hs;
  int device = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nsteps = -1;
  std::vector<double> taus;
  std::vector<int> versions;
  std::vector<hipEvent_t> events;
  hipEvent_t done = nullptr;  // recorded after each launch; every member's stream waits on it
  std::vector<hipEvent_t> pre;  // each member's stream before a launch: the launch waits on them
  void drop() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    nsteps = -1;
  }
  ~festung_batch() {
    drop();
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto e : pre) (void)hipEventDestroy(e);
    if (done) (void)hipEventDestroy(done);
  }
};

extern "C" {

int festung_batch_create(festung_problem* const* hs, int32_t n, festung_batch** out) {
  return guard([&] {
    if (!hs || !out || n < 1) fail(FESTUNG_ERR_ARG, "invalid argument");
    *out = nullptr;
    std::unique_ptr<festung_batch> b(new festung_batch);
    for (int32_t i = 0; i < n; ++i) {
      check_handle(hs[i]);
      if (hs[i]->device != hs[0]->device) fail(FESTUNG_ERR_ARG, "batched problems must share one device");
      for (int32_t j = 0; j < i; ++j)
        if (hs[j] == hs[i]) fail(FESTUNG_ERR_ARG, "problem %d appears twice", i);
      b->hs.push_back(hs[i]);
    }
    hs[0]->set_device();
    b->device = hs[0]->device;
    b->events.resize(size_t(n) + 1);
    for (auto& e : b->events) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->pre.resize(size_t(n));
    for (auto& e : b->pre) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    *out = b.release();
  });
}

void festung_batch_destroy(festung_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipDeviceSynchronize();  // the graph may still run; problems may already be gone
  delete b;
}

int festung_batch_step(festung_batch* b, int32_t nsteps, const double* tau) {
  return guard([&] {
    if (!b || !tau || nsteps < 0) fail(FESTUNG_ERR_ARG, "invalid argument");
    if (nsteps == 0) return;
    festung_problem* h0 = b->hs[0];
    h0->set_device();
    const size_t n = b->hs.size();
    for (auto* h : b->hs) {
      h->require_velocity();
      h->finish_stage_sequence();
      if (h->prof) fail(FESTUNG_ERR_STATE, "profiling is on for a batched problem");
    }
    std::vector<double> taus(tau, tau + n);
    std::vector<int> versions;
    for (auto* h : b->hs) versions.push_back(h->version);
    if (!b->exec || b->nsteps != nsteps || b->taus != taus || b->versions != versions) {
      if (b->exec) HIPCHK(hipStreamSynchronize(h0->stream));  // the last launch may still run
      b->drop();
      HIPCHK(hipStreamBeginCapture(h0->stream, hipStreamCaptureModeThreadLocal));
      try {
        HIPCHK(hipEventRecord(b->events[0], h0->stream));
        for (size_t i = 1; i < n; ++i) HIPCHK(hipStreamWaitEvent(b->hs[i]->stream, b->events[0], 0));
        for (size_t i = 0; i < n; ++i)
          for (int s = 0; s < nsteps; ++s) b->hs[i]->enqueue_step(taus[i]);
        for (size_t i = 1; i < n; ++i) {
          HIPCHK(hipEventRecord(b->events[i], b->hs[i]->stream));
          HIPCHK(hipStreamWaitEvent(h0->stream, b->events[i], 0));
        }
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(h0->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHK(hipStreamEndCapture(h0->stream, &b->graph));
      HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->nsteps = nsteps;
      b->taus = taus;
      b->versions = versions;
    }
    // work already enqueued on a member's own stream (set_state, a
    // coefficient change) runs before the batch
    for (size_t i = 1; i < n; ++i) {
      HIPCHK(hipEventRecord(b->pre[i], b->hs[i]->stream));
      HIPCHK(hipStreamWaitEvent(h0->stream, b->pre[i], 0));
    }
    HIPCHK(hipGraphLaunch(b->exec, h0->stream));
    // later work on any member's own stream (its per-problem entries) runs
    // after the batch, without a host synchronisation
    HIPCHK(hipEventRecord(b->done, h0->stream));
    for (size_t i = 1; i < n; ++i) HIPCHK(hipStreamWaitEvent(b->hs[i]->stream, b->done, 0));
    for (auto* h : b->hs) h->cur = 0;
  });
}

int festung_batch_sync(festung_batch* b) {
  return guard([&] {
    if (!b) fail(FESTUNG_ERR_ARG, "null batch");
    b->hs[0]->set_device();
    HIPCHK(hipStreamSynchronize(b->hs[0]->stream));
  });
}

int festung_sync(festung_problem* h) {
  return guard([&] {
    check_handle(h);
    h->set_device();
    sync(h);
  });
}

}  // extern "C"
