// Host-only part of the C-ABI: error reporting, grids and reference tables.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>

#include "checkpoint.hpp"
#include "common.hpp"
#include "grid.hpp"
#include "tables.hpp"

namespace festung {
static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }
const char* last_error() { return g_last_error.c_str(); }

// ------------------------------------------------------------- tuning --
namespace {
using TuneTuple = std::tuple<std::string, int, int, int, int>;
TuneTuple tuple_of(const TuneKey& k) { return TuneTuple(k.device, k.p, k.ordRK, k.lim, k.lgK); }
std::mutex g_tuneMu;
std::map<TuneTuple, int> g_tune;

bool valid_key(int p, int ordRK, int lim, int lgK, int bs) {
  return p >= 0 && p <= 4 && ordRK >= 1 && ordRK <= 4 && lim >= -1 && lim <= 2 && lgK >= 0 && lgK <= 40 &&
         (bs == 64 || bs == 128 || bs == 256);
}

// Parse a tuning file into `out` (all or nothing).
void parse_tune_file(const char* path, std::map<TuneTuple, int>& out) {
  std::ifstream f(path);
  if (!f) fail(FESTUNG_ERR_ARG, "cannot open tuning file %s", path);
  std::string line;
  int no = 0;
  while (std::getline(f, line)) {
    ++no;
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream is(line);
    std::string dev;
    if (!(is >> dev)) continue;  // blank or comment
    int p, ordRK, lim, lgK, bs;
    std::string extra;
    if (!(is >> p >> ordRK >> lim >> lgK >> bs) || (is >> extra) || !valid_key(p, ordRK, lim, lgK, bs))
      fail(FESTUNG_ERR_ARG, "%s:%d: expected 'device p ordRK limiter log2K elemsPerBlock'", path, no);
    out[TuneTuple(dev, p, ordRK, lim, lgK)] = bs;
  }
}
}  // namespace

int tune_lookup(const TuneKey& k) {
  std::lock_guard<std::mutex> lk(g_tuneMu);
  const auto it = g_tune.find(tuple_of(k));
  return it == g_tune.end() ? 0 : it->second;
}

void tune_record(const TuneKey& k, int bs) {
  std::lock_guard<std::mutex> lk(g_tuneMu);
  g_tune[tuple_of(k)] = bs;
}

void tune_load_env() {
  static std::once_flag once;
  std::call_once(once, [] {
    const char* path = std::getenv("FESTUNG_TUNE_FILE");
    if (!path || !*path) return;
    std::ifstream probe(path);
    if (!probe) return;  // not written yet: nothing tuned
    std::map<TuneTuple, int> m;
    parse_tune_file(path, m);
    std::lock_guard<std::mutex> lk(g_tuneMu);
    for (const auto& e : m) g_tune[e.first] = e.second;
  });
}
}  // namespace festung

using namespace festung;

namespace {
template <class T>
std::vector<int32_t> widen(const std::vector<T>& v) {
  return std::vector<int32_t>(v.begin(), v.end());
}

bool field_i32(const Grid& g, const std::string& f, std::vector<int32_t>& out) {
  if (f == "V0T") out = g.V0T;
  else if (f == "V0E") out = g.V0E;
  else if (f == "T0E") out = g.T0E;
  else if (f == "E0E") out = g.E0E;
  else if (f == "E0T") out = g.E0T;
  else if (f == "idE") out = g.idE;
  else if (f == "idE0T") out = g.idE0T;
  else if (f == "nbr") out = g.nbr;
  else if (f == "nbrEdge") out = g.nbrEdge;
  else if (f == "v2tPtr") out = g.v2tPtr;
  else if (f == "v2tIdx") out = g.v2tIdx;
  else if (f == "markE0Tint") out = widen(g.markE0Tint);
  else if (f == "markE0TbdrD") out = widen(g.markE0TbdrD);
  else if (f == "markE0TbdrN") out = widen(g.markE0TbdrN);
  else if (f == "markV0TbdrD") out = widen(g.markV0TbdrD);
  else return false;
  return true;
}

const std::vector<double>* field_f64(const Grid& g, const std::string& f) {
  if (f == "coordV") return &g.coordV;
  if (f == "areaE") return &g.areaE;
  if (f == "nuE") return &g.nuE;
  if (f == "baryE") return &g.baryE;
  if (f == "areaT") return &g.areaT;
  if (f == "baryT") return &g.baryT;
  if (f == "B") return &g.B;
  if (f == "areaE0T") return &g.areaE0T;
  if (f == "sigE0T") return &g.sigE0T;
  if (f == "nuE0T") return &g.nuE0T;
  if (f == "coordV0T") return &g.coordV0T;
  return nullptr;
}
}  // namespace

namespace {
std::vector<double> ref_table(int p, const std::string& name, int ord) {
  if (name == "quad1D") {
    Rule1D r = quad_rule_1d(ord);
    std::vector<double> v(r.q);
    v.insert(v.end(), r.w.begin(), r.w.end());
    return v;
  }
  if (name == "quad2D") {
    Rule2D r = quad_rule_2d(ord);
    std::vector<double> v(r.q1);
    v.insert(v.end(), r.q2.begin(), r.q2.end());
    v.insert(v.end(), r.w.begin(), r.w.end());
    return v;
  }
  RefTables t = build_ref_tables(p);
  const int N = t.N;
  if (name == "hatM") return t.hatM;
  if (name == "hatG") return t.hatG;
  if (name == "quadG") {
    std::vector<double> v(t.ruleG.q1);
    v.insert(v.end(), t.ruleG.q2.begin(), t.ruleG.q2.end());
    v.insert(v.end(), t.ruleG.w.begin(), t.ruleG.w.end());
    return v;
  }
  if (name == "hatRdiag") return t.hatRdiag;
  if (name == "hatRoff") return t.hatRoff;
  if (name == "taylorF") return t.F;
  if (name == "taylorFinv") return t.Finv;
  if (name == "phi2D" || name == "gradPhi2D") {
    Rule2D r = quad_rule_2d(ord);
    const int R = int(r.w.size());
    std::vector<double> v(size_t(R) * N * (name == "phi2D" ? 1 : 2));
    for (int i = 0; i < N; ++i)
      for (int q = 0; q < R; ++q) {
        if (name == "phi2D") v[q + i * R] = phi(i + 1, r.q1[q], r.q2[q]);
        else
          for (int m = 0; m < 2; ++m) v[q + i * R + m * R * N] = grad_phi(i + 1, m + 1, r.q1[q], r.q2[q]);
      }
    return v;
  }
  if (name == "phi1D" || name == "thetaPhi1D") {
    Rule1D r = quad_rule_1d(ord);
    const int R = int(r.w.size());
    std::vector<double> v(size_t(R) * N * (name == "phi1D" ? 3 : 9));
    for (int nn = 0; nn < 3; ++nn)
      for (int q = 0; q < R; ++q) {
        double a1, a2;
        gamma_map(nn + 1, r.q[q], a1, a2);
        for (int i = 0; i < N; ++i) {
          if (name == "phi1D") v[q + i * R + nn * R * N] = phi(i + 1, a1, a2);
          else
            for (int np = 0; np < 3; ++np) {
              double b1, b2;
              theta_map(nn + 1, np + 1, a1, a2, b1, b2);
              v[q + i * R + nn * R * N + np * R * N * 3] = phi(i + 1, b1, b2);
            }
        }
      }
    return v;
  }
  fail(FESTUNG_ERR_ARG, "unknown reference table '%s'", name.c_str());
}
}  // namespace

extern "C" {

const char* festung_last_error(void) { return g_last_error.c_str(); }
const char* festung_version(void) { return "festung-hip 0.1 (gfx950, fp64)"; }

int festung_domain_square(double hmax, double a, double b, festung_grid** out) {
  return guard([&] {
    if (!out) fail(FESTUNG_ERR_ARG, "null output pointer");
    *out = nullptr;
    auto* h = new festung_grid;
    try {
      h->g = domain_square(hmax, a, b);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int festung_grid_create(const double* coordV, int32_t numV, const int32_t* V0T, int32_t numT, festung_grid** out) {
  return guard([&] {
    if (!out || !coordV || !V0T) fail(FESTUNG_ERR_ARG, "null pointer argument");
    *out = nullptr;
    auto* h = new festung_grid;
    try {
      h->g = generate_grid_data(coordV, numV, V0T, numT);
      mark_boundaries(h->g, nullptr);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int festung_grid_set_boundary(festung_grid* h, const int32_t* idE, const uint8_t* markN) {
  return guard([&] {
    if (!h || !idE) fail(FESTUNG_ERR_ARG, "null pointer argument");
    Grid& g = h->g;
    g.idE.assign(idE, idE + g.numE);
    for (size_t i = 0; i < g.E0T.size(); ++i) g.idE0T[i] = g.idE[g.E0T[i]];
    mark_boundaries(g, markN);
  });
}

int festung_grid_concat(const festung_grid* const* grids, int32_t n, festung_grid** out) {
  return guard([&] {
    if (!grids || !out || n < 1) fail(FESTUNG_ERR_ARG, "invalid argument");
    *out = nullptr;
    std::vector<const Grid*> gs;
    for (int32_t i = 0; i < n; ++i) {
      if (!grids[i]) fail(FESTUNG_ERR_ARG, "null grid %d", i);
      gs.push_back(&grids[i]->g);
    }
    auto* h = new festung_grid;
    try {
      h->g = concat_grids(gs);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int festung_grid_sizes(const festung_grid* h, int32_t* numT, int32_t* numV, int32_t* numE) {
  return guard([&] {
    if (!h) fail(FESTUNG_ERR_ARG, "null grid");
    if (numT) *numT = h->g.numT;
    if (numV) *numV = h->g.numV;
    if (numE) *numE = h->g.numE;
  });
}


int festung_grid_field_size(const festung_grid* h, const char* field, int64_t* count, int* is_double) {
  return guard([&] {
    if (!h || !field || !count) fail(FESTUNG_ERR_ARG, "null pointer argument");
    std::vector<int32_t> vi;
    if (field_i32(h->g, field, vi)) {
      *count = int64_t(vi.size());
      if (is_double) *is_double = 0;
      return;
    }
    if (auto* v = field_f64(h->g, field)) {
      *count = int64_t(v->size());
      if (is_double) *is_double = 1;
      return;
    }
    fail(FESTUNG_ERR_ARG, "unknown grid field '%s'", field);
  });
}

int festung_grid_get_i32(const festung_grid* h, const char* field, int32_t* out) {
  return guard([&] {
    if (!h || !field || !out) fail(FESTUNG_ERR_ARG, "null pointer argument");
    std::vector<int32_t> vi;
    if (!field_i32(h->g, field, vi)) fail(FESTUNG_ERR_ARG, "unknown int32 grid field '%s'", field);
    std::memcpy(out, vi.data(), vi.size() * sizeof(int32_t));
  });
}

int festung_grid_get_f64(const festung_grid* h, const char* field, double* out) {
  return guard([&] {
    if (!h || !field || !out) fail(FESTUNG_ERR_ARG, "null pointer argument");
    auto* v = field_f64(h->g, field);
    if (!v) fail(FESTUNG_ERR_ARG, "unknown double grid field '%s'", field);
    std::memcpy(out, v->data(), v->size() * sizeof(double));
  });
}

int festung_grid_map_ref_to_phy(const festung_grid* h, int32_t nq, const double* q1, const double* q2, double* X1,
                                double* X2) {
  return guard([&] {
    if (!h || !q1 || !q2 || !X1 || !X2 || nq < 0) fail(FESTUNG_ERR_ARG, "invalid argument");
    map_ref_to_phy(h->g, nq, q1, q2, X1, X2);
  });
}

void festung_grid_destroy(festung_grid* h) { delete h; }


int festung_ref_size(int p, const char* name, int ord, int64_t* count) {
  return guard([&] {
    if (!name || !count) fail(FESTUNG_ERR_ARG, "null pointer argument");
    *count = int64_t(ref_table(p, name, ord).size());
  });
}

int festung_phi_at_points(int p, int32_t npts, const double* X1, const double* X2, double* out) {
  return guard([&] {
    if (p < 0 || p > 4) fail(FESTUNG_ERR_UNSUPPORTED, "polynomial degree p = %d (must be 0..4)", p);
    if (npts < 0 || (npts > 0 && (!X1 || !X2 || !out))) fail(FESTUNG_ERR_ARG, "invalid argument");
    const int N = nbasis(p);
    for (int i = 0; i < N; ++i)
      for (int32_t r = 0; r < npts; ++r) out[r + size_t(i) * npts] = phi(i + 1, X1[r], X2[r]);
  });
}

int festung_ref_get(int p, const char* name, int ord, double* out) {
  return guard([&] {
    if (!name || !out) fail(FESTUNG_ERR_ARG, "null pointer argument");
    auto v = ref_table(p, name, ord);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

}  // extern "C"

int festung_binomial_schedule(int32_t nsteps, int32_t snapshots, int64_t* advances, int64_t* actions) {
  return guard([&] {
    if (nsteps < 1 || snapshots < 1 || !advances || !actions) fail(FESTUNG_ERR_ARG, "invalid argument");
    std::vector<CkAction> plan;
    try {
      plan = binomial_schedule(nsteps, snapshots);
    } catch (const std::exception& e) {
      fail(FESTUNG_ERR_INTERNAL, "%s", e.what());
    }
    int64_t adv = 0;
    for (const CkAction& x : plan)
      if (x.kind == CkAction::ADVANCE) adv += x.steps;
    *advances = adv;
    *actions = int64_t(plan.size());
  });
}

int festung_tune_save(const char* path) {
  return guard([&] {
    if (!path) fail(FESTUNG_ERR_ARG, "null pointer argument");
    std::ostringstream os;
    os << "# festung tuning: device p ordRK limiter(-1 none, 0 linear, 1 hierarch_vert, 2 strict) "
          "log2K elemsPerBlock\n";
    {
      std::lock_guard<std::mutex> lk(g_tuneMu);
      for (const auto& e : g_tune)
        os << std::get<0>(e.first) << ' ' << std::get<1>(e.first) << ' ' << std::get<2>(e.first) << ' '
           << std::get<3>(e.first) << ' ' << std::get<4>(e.first) << ' ' << e.second << '\n';
    }
    std::ofstream f(path, std::ios::trunc);
    if (!f) fail(FESTUNG_ERR_ARG, "cannot write tuning file %s", path);
    f << os.str();
    if (!f) fail(FESTUNG_ERR_ARG, "cannot write tuning file %s", path);
  });
}

int festung_tune_load(const char* path) {
  return guard([&] {
    if (!path) fail(FESTUNG_ERR_ARG, "null pointer argument");
    std::map<TuneTuple, int> m;
    parse_tune_file(path, m);
    std::lock_guard<std::mutex> lk(g_tuneMu);
    for (const auto& e : m) g_tune[e.first] = e.second;
  });
}

int festung_tune_clear(void) {
  return guard([&] {
    std::lock_guard<std::mutex> lk(g_tuneMu);
    g_tune.clear();
  });
}

int festung_tune_count(int32_t* n) {
  return guard([&] {
    if (!n) fail(FESTUNG_ERR_ARG, "null pointer argument");
    std::lock_guard<std::mutex> lk(g_tuneMu);
    *n = int32_t(g_tune.size());
  });
}
