// festung_mex.cpp -- Matlab / GNU Octave MEX gateway over the C-ABI of
// libfestung_hip.so (include/festung_hip.h).
//
//   mex -R2018a -I<repo>/include festung_mex.cpp -L<repo>/festung_amd -lfestung_hip
//   (Octave: mkoctfile --mex -I<repo>/include festung_mex.cpp -L... -lfestung_hip)
//
// One MEX function dispatching on a command string, festung_mex(cmd, ...).
// Grid and problem handles are uint64 scalars; the gateway keeps the set of
// live handles (mexLock while any exist, freed by mexAtExit) and rejects
// handles it did not create.  Arrays keep the reference's shapes and
// column-major layouts, so coefficient matrices pass without copies; ids are
// converted between Matlab's 1-based doubles (0 = none) and the ABI's 0-based
// int32 (-1 = none).  Errors are raised with mexErrMsgIdAndTxt after every
// scratch buffer of the call has been released (it does not return).
//
// Commands (K x N arrays are column-major; x may also be the KN x 1
// element-major vector cDiscRK{s} of advection/solveStep.m:68):
//   g = festung_mex('grid_square', hmax)                     core/domainSquare.m
//   g = festung_mex('grid_create', coordV, V0T[, idE[, markE0TbdrN]])  core/generateGridData.m
//   v = festung_mex('grid_field', g, name)                   fields of the struct g
//   [X1, X2] = festung_mex('map_ref_to_phy', g, Q1, Q2)      g.mapRef2Phy
//   festung_mex('grid_destroy', g)
//   h = festung_mex('create', g, p, ordRK, flags)            advection/preprocessProblem.m
//   [K, N, RE, RF, Re] = festung_mex('sizes', h)
//   festung_mex('destroy', h)
//   festung_mex('velocity', h, u1Disc, u2Disc, vNormalOnQuadEdge)   ([] keeps a field)
//   festung_mex('velocity_points', h, u1q, u2q, u1e, u2e)
//   festung_mex('source', h, fq, cDe, gNe)                   ([] for zero data)
//   festung_mex('vertex_dirichlet', h, vD)                   ([] clears)
//   c = festung_mex('project', h, ord, fq[, setState])
//   festung_mex('set_state', h, cDisc);  c = festung_mex('get_state', h)
//   festung_mex('begin_step', h); festung_mex('stage', h, nSubStep, tau)
//   festung_mex('step', h, nsteps, tau)
//   cDot = festung_mex('rhs', h)
//   y = festung_mex('apply', h, op, x)    op: 'G1' 'G2' 'G' 'R' 'M' 'Minv'
//   [L, KD, KN] = festung_mex('vectors', h, fq, cDe, gNe)
//   [alphaE, minV0T, maxV0T] = festung_mex('vertex_limiter', h, valCentroid, valV0T, markV0TbdrD, dataV0T)
//   [dataDisc, minV0T, maxV0T] = festung_mex('slope_limiter', h, dataDisc, markV0TbdrD, dataV0T)
//   [alphaE, minV0T, maxV0T] = festung_mex('limit' | 'limit_disc_linear', h)
//   err = festung_mex('l2_error', h, ord, cq)
//   [t, omega] = festung_mex('rk', ord, tau, t0)             core/rungeKuttaExplicit.m
//   T = festung_mex('ref', p, name, ord)                     reference tables
//   festung_mex('sync', h);  s = festung_mex('version')
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "festung_hip.h"
#include "mex.h"

namespace {

struct GatewayError : std::runtime_error {
  std::string id;
  GatewayError(const std::string& i, const std::string& m) : std::runtime_error(m), id(i) {}
};

std::set<festung_grid*> g_grids;
std::map<festung_problem*, int32_t> g_problems;  // live problems and their vertex counts
bool g_atexit = false;

void release_all() {
  for (auto& hv : g_problems) festung_problem_destroy(hv.first);
  for (auto* g : g_grids) festung_grid_destroy(g);
  g_problems.clear();
  g_grids.clear();
}

void relock() {
  // locked while any handle lives, so that `clear festung_mex` cannot drop them
  static bool locked = false;
  const bool want = !g_problems.empty() || !g_grids.empty();
  if (want && !locked) mexLock();
  if (!want && locked) mexUnlock();
  locked = want;
}

void check(int status) {
  if (status != FESTUNG_OK) throw GatewayError("festung:hip", festung_last_error());
}

[[noreturn]] void bad(const std::string& m) { throw GatewayError("festung:arg", m); }

void need_args(int nrhs, int n, const char* usage) {
  if (nrhs < n) bad(std::string("usage: ") + usage);
}

std::string string_arg(const mxArray* a, const char* what) {
  if (!mxIsChar(a)) bad(std::string(what) + " must be a string");
  char buf[128];
  if (mxGetString(a, buf, sizeof buf) != 0) bad(std::string(what) + " too long");
  return buf;
}

double scalar_arg(const mxArray* a, const char* what) {
  if (!mxIsDouble(a) || mxGetNumberOfElements(a) != 1) bad(std::string(what) + " must be a double scalar");
  return mxGetScalar(a);
}

uint64_t handle_bits(const mxArray* a) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) bad("expected a festung handle (uint64 scalar)");
  return *static_cast<const uint64_t*>(mxGetData(a));
}

festung_problem* problem_arg(const mxArray* a) {
  auto* h = reinterpret_cast<festung_problem*>(static_cast<uintptr_t>(handle_bits(a)));
  if (!g_problems.count(h)) bad("not a live problem handle");
  return h;
}

festung_grid* grid_arg(const mxArray* a) {
  auto* g = reinterpret_cast<festung_grid*>(static_cast<uintptr_t>(handle_bits(a)));
  if (!g_grids.count(g)) bad("not a live grid handle");
  return g;
}

mxArray* handle_out(const void* p) {
  mxArray* a = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *static_cast<uint64_t*>(mxGetData(a)) = static_cast<uint64_t>(reinterpret_cast<uintptr_t>(p));
  return a;
}

bool is_empty(const mxArray* a) { return mxGetNumberOfElements(a) == 0; }

// a real double array with exactly n values ([] -> null when allowed)
const double* doubles(const mxArray* a, size_t n, const char* what, bool allowEmpty = false) {
  if (allowEmpty && is_empty(a)) return nullptr;
  if (!mxIsDouble(a) || mxIsComplex(a)) bad(std::string(what) + " must be a real double array");
  if (mxGetNumberOfElements(a) != n)
    bad(std::string(what) + " has " + std::to_string(mxGetNumberOfElements(a)) + " values, expected " +
        std::to_string(n));
  return mxGetPr(a);
}

// logical (or double 0/1) array of n values as bytes
std::vector<uint8_t> marks(const mxArray* a, size_t n, const char* what) {
  if (mxGetNumberOfElements(a) != n) bad(std::string(what) + " has the wrong size");
  std::vector<uint8_t> m(n);
  if (mxIsLogical(a)) {
    const mxLogical* l = mxGetLogicals(a);
    for (size_t i = 0; i < n; ++i) m[i] = l[i] ? 1 : 0;
  } else if (mxIsDouble(a)) {
    const double* d = mxGetPr(a);
    for (size_t i = 0; i < n; ++i) m[i] = d[i] != 0.0 ? 1 : 0;
  } else {
    bad(std::string(what) + " must be logical");
  }
  return m;
}

// 1-based Matlab ids (doubles, 0 = none) -> 0-based int32 (-1 = none)
std::vector<int32_t> ids_in(const mxArray* a, size_t n, const char* what) {
  const double* d = doubles(a, n, what);
  std::vector<int32_t> v(n);
  for (size_t i = 0; i < n; ++i) {
    if (d[i] != std::floor(d[i]) || d[i] < 0 || d[i] > 2147483647.0) bad(std::string(what) + " must hold ids");
    v[i] = static_cast<int32_t>(d[i]) - 1;
  }
  return v;
}

mxArray* doubles_out(size_t m, size_t n) { return mxCreateDoubleMatrix(m, n, mxREAL); }

struct Sizes {
  int32_t K, N, RE, RF, Re;
};
Sizes sizes(festung_problem* h) {
  Sizes s{};
  check(festung_problem_sizes(h, &s.K, &s.N, &s.RE, &s.RF, &s.Re));
  return s;
}

// A coefficient argument: K x N column-major, or the KN x 1 element-major
// vector of advection/solveStep.m:68 (converted; the result is returned in
// the same form).
struct Coeffs {
  std::vector<double> soa;
  bool elementMajor = false;
};
Coeffs coeffs_in(const mxArray* a, const Sizes& s, const char* what) {
  Coeffs c;
  const size_t kn = size_t(s.K) * s.N;
  const double* d = doubles(a, kn, what);
  c.elementMajor = s.N > 1 && mxGetN(a) == 1;
  c.soa.assign(d, d + kn);
  if (c.elementMajor)
    for (int32_t k = 0; k < s.K; ++k)
      for (int32_t i = 0; i < s.N; ++i) c.soa[size_t(i) * s.K + k] = d[size_t(k) * s.N + i];
  return c;
}
mxArray* coeffs_out(const std::vector<double>& soa, const Sizes& s, bool elementMajor) {
  const size_t kn = size_t(s.K) * s.N;
  mxArray* o = elementMajor ? doubles_out(kn, 1) : doubles_out(s.K, s.N);
  double* d = mxGetPr(o);
  if (!elementMajor) {
    std::memcpy(d, soa.data(), kn * sizeof(double));
  } else {
    for (int32_t k = 0; k < s.K; ++k)
      for (int32_t i = 0; i < s.N; ++i) d[size_t(k) * s.N + i] = soa[size_t(i) * s.K + k];
  }
  return o;
}

int op_code(const std::string& op) {
  if (op == "G") return FESTUNG_OP_G;
  if (op == "G1") return FESTUNG_OP_G1;
  if (op == "G2") return FESTUNG_OP_G2;
  if (op == "R") return FESTUNG_OP_R;
  if (op == "M") return FESTUNG_OP_M;
  if (op == "Minv") return FESTUNG_OP_MINV;
  bad("unknown operator '" + op + "' (G1, G2, G, R, M, Minv)");
}

const char* const kIdFields[] = {"V0T", "V0E", "T0E", "E0E", "E0T", "nbr", "v2tIdx"};
const char* const kMarkFields[] = {"markE0Tint", "markE0TbdrD", "markE0TbdrN", "markV0TbdrD"};

void grid_field(int nlhs, mxArray* plhs[], festung_grid* g, const std::string& name) {
  (void)nlhs;
  int32_t T, V, E;
  check(festung_grid_sizes(g, &T, &V, &E));
  int64_t cnt = 0;
  int isd = 0;
  check(festung_grid_field_size(g, name.c_str(), &cnt, &isd));
  // rows: the field's first dimension in the struct g (core/generateGridData.m)
  size_t rows = size_t(T);
  if (name == "coordV") rows = size_t(V);
  if (name == "V0E" || name == "T0E" || name == "E0E" || name == "idE" || name == "areaE" || name == "nuE" ||
      name == "baryE")
    rows = size_t(E);
  if (name == "v2tPtr" || name == "v2tIdx") rows = size_t(cnt);
  const size_t cols = rows ? size_t(cnt) / rows : 0;
  if (isd) {
    plhs[0] = doubles_out(rows, cols);
    check(festung_grid_get_f64(g, name.c_str(), mxGetPr(plhs[0])));
    return;
  }
  std::vector<int32_t> v(static_cast<size_t>(cnt));
  check(festung_grid_get_i32(g, name.c_str(), v.data()));
  bool isId = false, isMark = false;
  for (const char* f : kIdFields) isId |= name == f;
  for (const char* f : kMarkFields) isMark |= name == f;
  if (isMark) {
    plhs[0] = mxCreateLogicalMatrix(rows, cols);
    mxLogical* l = mxGetLogicals(plhs[0]);
    for (size_t i = 0; i < v.size(); ++i) l[i] = v[i] != 0;
    return;
  }
  plhs[0] = doubles_out(rows, cols);
  double* d = mxGetPr(plhs[0]);
  for (size_t i = 0; i < v.size(); ++i) d[i] = isId ? double(v[i] + 1) : double(v[i]);
}

void dispatch(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 1) bad("usage: festung_mex(command, ...)");
  const std::string cmd = string_arg(prhs[0], "command");
  if (cmd == "version") {
    plhs[0] = mxCreateString(festung_version());
  } else if (cmd == "grid_square") {
    need_args(nrhs, 2, "g = festung_mex('grid_square', hmax)");
    festung_grid* g = nullptr;
    check(festung_domain_square(scalar_arg(prhs[1], "hmax"), 0.0, 1.0, &g));
    g_grids.insert(g);
    plhs[0] = handle_out(g);
  } else if (cmd == "grid_create") {
    need_args(nrhs, 3, "g = festung_mex('grid_create', coordV, V0T[, idE[, markE0TbdrN]])");
    const size_t nV = mxGetM(prhs[1]), K = mxGetM(prhs[2]);
    if (mxGetN(prhs[1]) != 2 || mxGetN(prhs[2]) != 3) bad("coordV must be numV x 2 and V0T K x 3");
    const double* cv = doubles(prhs[1], 2 * nV, "coordV");
    std::vector<int32_t> v0t = ids_in(prhs[2], 3 * K, "V0T");
    for (int32_t id : v0t)
      if (id < 0 || size_t(id) >= nV) bad("V0T holds a vertex id out of range");
    festung_grid* g = nullptr;
    check(festung_grid_create(cv, int32_t(nV), v0t.data(), int32_t(K), &g));
    try {
      if (nrhs >= 4 && !is_empty(prhs[3])) {
        int32_t T, V, E;
        check(festung_grid_sizes(g, &T, &V, &E));
        const double* d = doubles(prhs[3], size_t(E), "idE");
        std::vector<int32_t> idE(d, d + E);
        std::vector<uint8_t> mN;
        if (nrhs >= 5 && !is_empty(prhs[4])) mN = marks(prhs[4], 3 * K, "markE0TbdrN");
        check(festung_grid_set_boundary(g, idE.data(), mN.empty() ? nullptr : mN.data()));
      }
    } catch (...) {
      festung_grid_destroy(g);
      throw;
    }
    g_grids.insert(g);
    plhs[0] = handle_out(g);
  } else if (cmd == "grid_field") {
    need_args(nrhs, 3, "v = festung_mex('grid_field', g, name)");
    grid_field(nlhs, plhs, grid_arg(prhs[1]), string_arg(prhs[2], "field name"));
  } else if (cmd == "map_ref_to_phy") {
    need_args(nrhs, 4, "[X1, X2] = festung_mex('map_ref_to_phy', g, Q1, Q2)");
    festung_grid* g = grid_arg(prhs[1]);
    const size_t nq = mxGetNumberOfElements(prhs[2]);
    const double* q1 = doubles(prhs[2], nq, "Q1");
    const double* q2 = doubles(prhs[3], nq, "Q2");
    int32_t T, V, E;
    check(festung_grid_sizes(g, &T, &V, &E));
    mxArray* x1 = doubles_out(size_t(T), nq);
    mxArray* x2 = doubles_out(size_t(T), nq);
    plhs[0] = x1;
    if (nlhs > 1) plhs[1] = x2;
    check(festung_grid_map_ref_to_phy(g, int32_t(nq), q1, q2, mxGetPr(x1), mxGetPr(x2)));
    if (nlhs <= 1) mxDestroyArray(x2);
  } else if (cmd == "grid_destroy") {
    need_args(nrhs, 2, "festung_mex('grid_destroy', g)");
    festung_grid* g = grid_arg(prhs[1]);
    g_grids.erase(g);
    festung_grid_destroy(g);
  } else if (cmd == "create") {
    need_args(nrhs, 5, "h = festung_mex('create', g, p, ordRK, flags)");
    festung_problem* h = nullptr;
    check(festung_problem_create(grid_arg(prhs[1]), int(scalar_arg(prhs[2], "p")), int(scalar_arg(prhs[3], "ordRK")),
                                 int(scalar_arg(prhs[4], "flags")), &h));
    int32_t T, V, E;
    check(festung_grid_sizes(grid_arg(prhs[1]), &T, &V, &E));
    g_problems[h] = V;
    plhs[0] = handle_out(h);
  } else if (cmd == "destroy") {
    need_args(nrhs, 2, "festung_mex('destroy', h)");
    festung_problem* h = problem_arg(prhs[1]);
    g_problems.erase(h);
    festung_problem_destroy(h);
  } else if (cmd == "rk") {
    need_args(nrhs, 4, "[t, omega] = festung_mex('rk', ord, tau, t0)");
    double t[4], w[4];
    int32_t S = 0;
    check(festung_runge_kutta_explicit(int(scalar_arg(prhs[1], "ord")), scalar_arg(prhs[2], "tau"),
                                       scalar_arg(prhs[3], "t0"), t, w, &S));
    plhs[0] = doubles_out(1, size_t(S));
    std::memcpy(mxGetPr(plhs[0]), t, S * sizeof(double));
    if (nlhs > 1) {
      plhs[1] = doubles_out(1, size_t(S));
      std::memcpy(mxGetPr(plhs[1]), w, S * sizeof(double));
    }
  } else if (cmd == "ref") {
    need_args(nrhs, 3, "T = festung_mex('ref', p, name[, ord])");
    const int p = int(scalar_arg(prhs[1], "p"));
    const std::string name = string_arg(prhs[2], "table name");
    const int ord = nrhs >= 4 ? int(scalar_arg(prhs[3], "ord")) : 0;
    int64_t cnt = 0;
    check(festung_ref_size(p, name.c_str(), ord, &cnt));
    plhs[0] = doubles_out(size_t(cnt), 1);
    check(festung_ref_get(p, name.c_str(), ord, mxGetPr(plhs[0])));
  } else {
    // commands on a problem handle
    static const std::set<std::string> kProblemCommands = {
        "sizes", "velocity", "velocity_points", "source", "vertex_dirichlet", "project", "set_state", "get_state",
        "begin_step", "stage", "step", "rhs", "apply", "vectors", "vertex_limiter", "slope_limiter", "limit",
        "limit_disc_linear", "l2_error", "sync"};
    if (!kProblemCommands.count(cmd)) bad("unknown command '" + cmd + "'");
    need_args(nrhs, 2, "festung_mex(command, h, ...)");
    festung_problem* h = problem_arg(prhs[1]);
    const Sizes s = sizes(h);
    const size_t K = size_t(s.K), KN = K * s.N, KE = 3 * K * size_t(s.Re), KQ = K * size_t(s.RE);
    if (cmd == "sizes") {
      const int32_t v[5] = {s.K, s.N, s.RE, s.RF, s.Re};
      for (int i = 0; i < 5 && i < (nlhs > 0 ? nlhs : 1); ++i) plhs[i] = mxCreateDoubleScalar(v[i]);
    } else if (cmd == "velocity") {
      need_args(nrhs, 5, "festung_mex('velocity', h, u1Disc, u2Disc, vNormalOnQuadEdge)");
      check(festung_set_velocity_disc(h, doubles(prhs[2], KN, "u1Disc", true), doubles(prhs[3], KN, "u2Disc", true),
                                      doubles(prhs[4], KE, "vNormalOnQuadEdge", true)));
    } else if (cmd == "velocity_points") {
      need_args(nrhs, 6, "festung_mex('velocity_points', h, u1q, u2q, u1e, u2e)");
      check(festung_set_velocity_points(h, doubles(prhs[2], KQ, "u1q"), doubles(prhs[3], KQ, "u2q"),
                                        doubles(prhs[4], KE, "u1e"), doubles(prhs[5], KE, "u2e")));
    } else if (cmd == "source") {
      need_args(nrhs, 5, "festung_mex('source', h, fq, cDe, gNe)");
      check(festung_set_source_points(h, doubles(prhs[2], KQ, "fq", true), doubles(prhs[3], KE, "cDe", true),
                                      doubles(prhs[4], KE, "gNe", true)));
    } else if (cmd == "vertex_dirichlet") {
      need_args(nrhs, 3, "festung_mex('vertex_dirichlet', h, vD)");
      check(festung_set_vertex_dirichlet(h, doubles(prhs[2], size_t(g_problems[h]), "vD", true)));
    } else if (cmd == "project") {
      need_args(nrhs, 4, "c = festung_mex('project', h, ord, fq[, setState])");
      const int ord = int(scalar_arg(prhs[2], "ord"));
      const size_t R = mxGetNumberOfElements(prhs[3]) / (K ? K : 1);
      plhs[0] = doubles_out(K, size_t(s.N));
      check(festung_project(h, ord, doubles(prhs[3], K * R, "fq"), mxGetPr(plhs[0]),
                            nrhs >= 5 ? int(scalar_arg(prhs[4], "setState")) : 0));
    } else if (cmd == "set_state") {
      need_args(nrhs, 3, "festung_mex('set_state', h, cDisc)");
      const Coeffs c = coeffs_in(prhs[2], s, "cDisc");
      check(festung_set_state(h, c.soa.data()));
    } else if (cmd == "get_state") {
      plhs[0] = doubles_out(K, size_t(s.N));
      check(festung_get_state(h, mxGetPr(plhs[0])));
    } else if (cmd == "begin_step") {
      check(festung_begin_step(h));
    } else if (cmd == "stage") {
      need_args(nrhs, 4, "festung_mex('stage', h, nSubStep, tau)");
      check(festung_stage(h, int(scalar_arg(prhs[2], "nSubStep")), scalar_arg(prhs[3], "tau")));
    } else if (cmd == "step") {
      need_args(nrhs, 4, "festung_mex('step', h, nsteps, tau)");
      check(festung_step(h, int(scalar_arg(prhs[2], "nsteps")), scalar_arg(prhs[3], "tau")));
    } else if (cmd == "rhs") {
      plhs[0] = doubles_out(K, size_t(s.N));
      check(festung_rhs(h, mxGetPr(plhs[0])));
    } else if (cmd == "apply") {
      need_args(nrhs, 4, "y = festung_mex('apply', h, op, x)");
      const int op = op_code(string_arg(prhs[2], "op"));
      const Coeffs x = coeffs_in(prhs[3], s, "x");
      std::vector<double> y(KN);
      check(festung_apply_operator(h, op, x.soa.data(), y.data()));
      plhs[0] = coeffs_out(y, s, x.elementMajor);
    } else if (cmd == "vectors") {
      need_args(nrhs, 5, "[L, KD, KN] = festung_mex('vectors', h, fq, cDe, gNe)");
      const double* fq = doubles(prhs[2], KQ, "fq", true);
      const double* cDe = doubles(prhs[3], KE, "cDe", true);
      const double* gNe = doubles(prhs[4], KE, "gNe", true);
      // as element-major KN x 1 vectors, like globL, globKD, globKN
      std::vector<double> v[3];
      const double* in[3] = {fq, cDe, gNe};
      for (int i = 0; i < 3; ++i)
        if (in[i]) v[i].resize(KN);
      check(festung_assemble_vectors(h, fq, cDe, gNe, fq ? v[0].data() : nullptr, cDe ? v[1].data() : nullptr,
                                     gNe ? v[2].data() : nullptr));
      for (int i = 0; i < 3 && i < (nlhs > 0 ? nlhs : 1); ++i)
        plhs[i] = in[i] ? coeffs_out(v[i], s, true) : doubles_out(0, 0);
    } else if (cmd == "vertex_limiter") {
      need_args(nrhs, 6, "[alphaE, minV0T, maxV0T] = festung_mex('vertex_limiter', h, valCentroid, valV0T, "
                         "markV0TbdrD, dataV0T)");
      const double* vc = doubles(prhs[2], K, "valCentroid");
      const double* vv = doubles(prhs[3], 3 * K, "valV0T");
      std::vector<uint8_t> mk;
      const double* dv = nullptr;
      if (!is_empty(prhs[4])) {
        mk = marks(prhs[4], 3 * K, "markV0TbdrD");
        dv = doubles(prhs[5], 3 * K, "dataV0T");
      }
      mxArray* a = doubles_out(K, 1);
      mxArray* mn = doubles_out(K, 3);
      mxArray* mx = doubles_out(K, 3);
      plhs[0] = a;
      if (nlhs > 1) plhs[1] = mn;
      if (nlhs > 2) plhs[2] = mx;
      check(festung_vertex_based_limiter(h, vc, vv, mk.empty() ? nullptr : mk.data(), dv, mxGetPr(a), mxGetPr(mn),
                                         mxGetPr(mx)));
      if (nlhs <= 1) mxDestroyArray(mn);
      if (nlhs <= 2) mxDestroyArray(mx);
    } else if (cmd == "slope_limiter") {
      need_args(nrhs, 5, "[dataDisc, minV0T, maxV0T] = festung_mex('slope_limiter', h, dataDisc, markV0TbdrD, "
                         "dataV0T)");
      const Coeffs c = coeffs_in(prhs[2], s, "dataDisc");
      std::vector<uint8_t> mk;
      const double* dv = nullptr;
      if (!is_empty(prhs[3])) {
        mk = marks(prhs[3], 3 * K, "markV0TbdrD");
        dv = doubles(prhs[4], 3 * K, "dataV0T");
      }
      std::vector<double> out(KN);
      mxArray* mn = doubles_out(K, 3);
      mxArray* mx = doubles_out(K, 3);
      try {
        check(festung_apply_slope_limiter_disc(h, c.soa.data(), mk.empty() ? nullptr : mk.data(), dv, out.data(),
                                               mxGetPr(mn), mxGetPr(mx)));
      } catch (...) {
        mxDestroyArray(mn);
        mxDestroyArray(mx);
        throw;
      }
      plhs[0] = coeffs_out(out, s, c.elementMajor);
      if (nlhs > 1) plhs[1] = mn; else mxDestroyArray(mn);
      if (nlhs > 2) plhs[2] = mx; else mxDestroyArray(mx);
    } else if (cmd == "limit" || cmd == "limit_disc_linear") {
      check(cmd == "limit" ? festung_limit(h) : festung_limit_disc_linear(h));
      mxArray* a = doubles_out(K, 1);
      mxArray* mn = doubles_out(K, 3);
      mxArray* mx = doubles_out(K, 3);
      plhs[0] = a;
      if (nlhs > 1) plhs[1] = mn;
      if (nlhs > 2) plhs[2] = mx;
      check(festung_limiter_info(h, mxGetPr(a), mxGetPr(mn), mxGetPr(mx)));
      if (nlhs <= 1) mxDestroyArray(mn);
      if (nlhs <= 2) mxDestroyArray(mx);
    } else if (cmd == "l2_error") {
      need_args(nrhs, 4, "err = festung_mex('l2_error', h, ord, cq)");
      const size_t R = mxGetNumberOfElements(prhs[3]) / (K ? K : 1);
      double e = 0;
      check(festung_l2_error(h, int(scalar_arg(prhs[2], "ord")), doubles(prhs[3], K * R, "cq"), &e));
      plhs[0] = mxCreateDoubleScalar(e);
    } else if (cmd == "sync") {
      check(festung_sync(h));
    } else {
      bad("unknown command '" + cmd + "'");
    }
  }
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (!g_atexit) {
    mexAtExit(release_all);
    g_atexit = true;
  }
  std::string id, msg;
  try {
    dispatch(nlhs, plhs, nrhs, prhs);
  } catch (const GatewayError& e) {
    id = e.id;
    msg = e.what();
  } catch (const std::exception& e) {
    id = "festung:internal";
    msg = e.what();
  }
  relock();
  // every scratch vector of the call is gone: raise the error (no return)
  if (!id.empty()) mexErrMsgIdAndTxt(id.c_str(), "%s", msg.c_str());
}
