// Mesh generation and connectivity (host side), following
// core/domainSquare.m:67-105 and core/generateGridData.m:130-198 with the
// reference's numbering semantics:
//  * edge ids are the column-major order of find(triu(V2T + V2T')), i.e.
//    undirected edges sorted by (larger vertex id, smaller vertex id);
//  * V0E / T0E come from the lowest-numbered triangle containing the edge
//    (the reverse-order, last-write-wins assignments of :142-148);
//  * E0T(k, n) is the edge opposite local vertex n.
// The vertex patches of :187 (dense K x K) are stored as a vertex -> element
// CSR list instead, which has the same pattern at O(K) memory.
#include "grid.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

#include "common.hpp"

namespace festung {

namespace {

// Directed or undirected edge lookup keyed by the first vertex: for each
// vertex a, a short list of (b, value).
struct EdgeMap {
  std::vector<int32_t> ptr, key, val;
  void build(int32_t numV, const std::vector<int32_t>& a, const std::vector<int32_t>& b,
             const std::vector<int32_t>& v) {
    ptr.assign(size_t(numV) + 1, 0);
    for (int32_t x : a) ++ptr[x + 1];
    for (int32_t i = 0; i < numV; ++i) ptr[i + 1] += ptr[i];
    key.resize(a.size());
    val.resize(a.size());
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (size_t i = 0; i < a.size(); ++i) {
      key[pos[a[i]]] = b[i];
      val[pos[a[i]]++] = v[i];
    }
  }
  int32_t find(int32_t a, int32_t b) const {
    for (int32_t i = ptr[a]; i < ptr[a + 1]; ++i)
      if (key[i] == b) return val[i];
    return -1;
  }
};

}  // namespace

Grid generate_grid_data(const double* coordV, int32_t numV, const int32_t* V0T, int32_t K) {
  if (numV < 3 || K < 1) fail(FESTUNG_ERR_ARG, "generateGridData: need numV >= 3 and numT >= 1");
  Grid g;
  g.numT = K;
  g.numV = numV;
  g.coordV.assign(coordV, coordV + 2 * size_t(numV));
  g.V0T.assign(V0T, V0T + 3 * size_t(K));
  auto v0t = [&](int32_t k, int n) { return g.V0T[k + size_t(n) * K]; };
  for (size_t i = 0; i < g.V0T.size(); ++i)
    if (g.V0T[i] < 0 || g.V0T[i] >= numV) fail(FESTUNG_ERR_ARG, "generateGridData: V0T entry out of range");
  const double* x = g.coordV.data();
  const double* y = g.coordV.data() + numV;

  // areaT, :153-156 (the reference's shoelace expression order)
  g.areaT.resize(K);
  for (int32_t k = 0; k < K; ++k) {
    int32_t a = v0t(k, 0), b = v0t(k, 1), c = v0t(k, 2);
    g.areaT[k] = (x[a] * y[b] + x[b] * y[c] + x[c] * y[a] - x[a] * y[c] - x[b] * y[a] - x[c] * y[b]) / 2;
    if (!(g.areaT[k] > 0))
      fail(FESTUNG_ERR_ARG, "generateGridData: element %d is not counter-clockwise (area %g)", k, g.areaT[k]);
  }

  // V2T: directed edge a->b of triangle k (:136-137); must be unique
  std::vector<int32_t> da(3 * size_t(K)), db(3 * size_t(K)), dk(3 * size_t(K));
  for (int32_t k = 0; k < K; ++k)
    for (int n = 0; n < 3; ++n) {
      da[3 * size_t(k) + n] = v0t(k, n);
      db[3 * size_t(k) + n] = v0t(k, (n + 1) % 3);
      dk[3 * size_t(k) + n] = k;
    }
  EdgeMap V2T;
  V2T.build(numV, da, db, dk);
  for (int32_t a = 0; a < numV; ++a)
    for (int32_t i = V2T.ptr[a]; i < V2T.ptr[a + 1]; ++i)
      for (int32_t j = i + 1; j < V2T.ptr[a + 1]; ++j)
        if (V2T.key[i] == V2T.key[j])
          fail(FESTUNG_ERR_ARG, "generateGridData: directed edge (%d,%d) in two elements", a, V2T.key[i]);

  // edge numbering (:139-141): undirected pairs sorted by (max, min)
  std::vector<int64_t> und;
  und.reserve(3 * size_t(K));
  for (size_t i = 0; i < da.size(); ++i) {
    int64_t lo = std::min(da[i], db[i]), hi = std::max(da[i], db[i]);
    und.push_back(hi * int64_t(numV) + lo);
  }
  std::sort(und.begin(), und.end());
  und.erase(std::unique(und.begin(), und.end()), und.end());
  const int32_t numE = int32_t(und.size());
  g.numE = numE;
  std::vector<int32_t> ea(numE), eb(numE), ev(numE);
  for (int32_t e = 0; e < numE; ++e) {
    ea[e] = int32_t(und[e] % numV);   // smaller vertex
    eb[e] = int32_t(und[e] / numV);   // larger vertex
    ev[e] = e;
  }
  EdgeMap V2E;
  V2E.build(numV, ea, eb, ev);
  auto edgeOf = [&](int32_t a, int32_t b) { return a < b ? V2E.find(a, b) : V2E.find(b, a); };

  // V0E / T0E from the lowest-numbered triangle (:142-148)
  g.V0E.assign(2 * size_t(numE), -1);
  g.T0E.assign(2 * size_t(numE), -1);
  for (int32_t k = 0; k < K; ++k)
    for (int n = 0; n < 3; ++n) {
      int32_t a = v0t(k, n), b = v0t(k, (n + 1) % 3);
      int32_t e = edgeOf(a, b);
      if (g.V0E[e] != -1) continue;
      g.V0E[e] = a;
      g.V0E[e + numE] = b;
      g.T0E[e] = V2T.find(a, b);
      g.T0E[e + numE] = V2T.find(b, a);
    }
  // areaE, nuE (:150-152)
  g.areaE.resize(numE);
  g.nuE.resize(2 * size_t(numE));
  g.baryE.resize(2 * size_t(numE));
  for (int32_t e = 0; e < numE; ++e) {
    int32_t a = g.V0E[e], b = g.V0E[e + numE];
    double vx = x[b] - x[a], vy = y[b] - y[a];
    g.areaE[e] = std::sqrt(vx * vx + vy * vy);
    g.nuE[e] = vy / g.areaE[e];
    g.nuE[e + numE] = -vx / g.areaE[e];
    g.baryE[e] = 0.5 * (x[a] + x[b]);
    g.baryE[e + numE] = 0.5 * (y[a] + y[b]);
  }
  // baryT (:157)
  g.baryT.resize(2 * size_t(K));
  for (int32_t k = 0; k < K; ++k) {
    int32_t a = v0t(k, 0), b = v0t(k, 1), c = v0t(k, 2);
    g.baryT[k] = (x[a] + x[b] + x[c]) / 3;
    g.baryT[k + K] = (y[a] + y[b] + y[c]) / 3;
  }
  // E0T (:158), areaE0T (:159), sigE0T (:160), coordV0T / nuE0T / E0E (:162-170)
  g.E0T.resize(3 * size_t(K));
  g.areaE0T.resize(3 * size_t(K));
  g.sigE0T.resize(3 * size_t(K));
  g.nuE0T.resize(6 * size_t(K));
  g.coordV0T.resize(6 * size_t(K));
  g.E0E.assign(2 * size_t(numE), -1);
  for (int32_t k = 0; k < K; ++k)
    for (int n = 0; n < 3; ++n) {
      int32_t e = edgeOf(v0t(k, (n + 1) % 3), v0t(k, (n + 2) % 3));
      size_t kn = k + size_t(n) * K;
      g.E0T[kn] = e;
      g.areaE0T[kn] = g.areaE[e];
      double sig = 1 - 2 * double(g.T0E[e + numE] == k);
      g.sigE0T[kn] = sig;
      for (int m = 0; m < 2; ++m) {
        g.coordV0T[kn + size_t(m) * 3 * K] = g.coordV[v0t(k, n) + size_t(m) * numV];
        g.nuE0T[kn + size_t(m) * 3 * K] = g.nuE[e + size_t(m) * numE] * sig;
      }
      if (sig == 1) g.E0E[e] = n;
      else g.E0E[e + numE] = n;
    }
  // B (:171-174): B(k, m, 1) = a2 - a1, B(k, m, 2) = a3 - a1
  g.B.resize(4 * size_t(K));
  for (int32_t k = 0; k < K; ++k)
    for (int m = 0; m < 2; ++m) {
      double a1 = g.coordV0T[k + size_t(m) * 3 * K];
      g.B[k + size_t(m) * K] = g.coordV0T[k + size_t(K) + size_t(m) * 3 * K] - a1;
      g.B[k + size_t(m) * K + 2 * size_t(K)] = g.coordV0T[k + 2 * size_t(K) + size_t(m) * 3 * K] - a1;
    }
  // neighbours (markE0TE0T{nn,np}(k, k') of :177-185)
  g.nbr.assign(3 * size_t(K), -1);
  g.nbrEdge.assign(3 * size_t(K), -1);
  for (int32_t k = 0; k < K; ++k)
    for (int n = 0; n < 3; ++n) {
      int32_t e = g.E0T[k + size_t(n) * K];
      if (g.E0E[e + numE] == -1) continue;  // boundary edge
      bool first = g.T0E[e] == k;
      g.nbr[k + size_t(n) * K] = first ? g.T0E[e + numE] : g.T0E[e];
      g.nbrEdge[k + size_t(n) * K] = first ? g.E0E[e + numE] : g.E0E[e];
    }
  // vertex -> element patches (CSR, ascending element ids)
  g.v2tPtr.assign(size_t(numV) + 1, 0);
  for (size_t i = 0; i < g.V0T.size(); ++i) ++g.v2tPtr[g.V0T[i] + 1];
  for (int32_t v = 0; v < numV; ++v) g.v2tPtr[v + 1] += g.v2tPtr[v];
  g.v2tIdx.resize(g.V0T.size());
  {
    std::vector<int32_t> pos(g.v2tPtr.begin(), g.v2tPtr.end() - 1);
    for (int32_t k = 0; k < K; ++k)
      for (int n = 0; n < 3; ++n) g.v2tIdx[pos[v0t(k, n)]++] = k;
  }
  g.idE.assign(numE, 0);
  g.idE0T.assign(3 * size_t(K), 0);
  return g;
}

static std::vector<double> matlab_colon(double a, double h, double b) {
  // a:h:b for h > 0, computed symmetrically (a + k h on the first half,
  // b - (n - k) h on the second half); exact for dyadic h.
  int n = int(std::floor((b - a) / h + 1e-10));
  std::vector<double> xs(n + 1);
  int half = (n + 1) / 2;
  for (int k = 0; k <= n; ++k) xs[k] = k < half ? a + k * h : b - (n - k) * h;
  return xs;
}

Grid domain_square(double h, double a, double b) {
  if (!(h > 0) || !(b > a)) fail(FESTUNG_ERR_ARG, "domainSquare: need h > 0 and b > a");
  double dimd = std::ceil((b - a) / h);
  if (dimd > 46000) fail(FESTUNG_ERR_ARG, "domainSquare: mesh too large (dim %g)", dimd);
  int dim = int(dimd);
  h = (b - a) / dim;
  std::vector<double> xs = matlab_colon(a, h, b);
  const int32_t nv1 = dim + 1, numV = nv1 * nv1;
  std::vector<double> coordV(2 * size_t(numV));
  for (int j = 0; j < nv1; ++j)      // x index (meshgrid column)
    for (int i = 0; i < nv1; ++i) {  // y index (meshgrid row), fastest
      coordV[i + size_t(j) * nv1] = xs[j];
      coordV[i + size_t(j) * nv1 + numV] = xs[i];
    }
  const int32_t K = 2 * dim * dim;
  std::vector<int32_t> V0T(3 * size_t(K));
  int32_t k = 0;
  const int32_t pat1[3] = {0, dim + 1, 1}, pat2[3] = {dim + 1, dim + 2, 1};
  for (int pass = 0; pass < 2; ++pass) {
    const int32_t* pat = pass == 0 ? pat1 : pat2;
    for (int32_t s = 0; s < dim * (dim + 1); ++s) {
      if (s % (dim + 1) == dim) continue;  // rows (dim+1):(dim+1):dim(dim+1) removed
      for (int n = 0; n < 3; ++n) V0T[k + size_t(n) * K] = pat[n] + s;
      ++k;
    }
  }
  Grid g = generate_grid_data(coordV.data(), numV, V0T.data(), K);
  g.dim = dim;
  const double tol = 10 * DBL_EPSILON;
  for (int32_t e = 0; e < g.numE; ++e) {
    double bx = g.baryE[e], by = g.baryE[e + g.numE];
    int id = 0;
    if (std::fabs(by - a) < tol) id = 1;
    if (std::fabs(bx - b) < tol) id = 2;
    if (std::fabs(by - b) < tol) id = 3;
    if (std::fabs(bx - a) < tol) id = 4;
    g.idE[e] = id;
  }
  for (size_t i = 0; i < g.E0T.size(); ++i) g.idE0T[i] = g.idE[g.E0T[i]];
  mark_boundaries(g, nullptr);
  return g;
}

void mark_boundaries(Grid& g, const uint8_t* markN) {
  const size_t n3 = 3 * size_t(g.numT);
  g.markE0Tint.resize(n3);
  g.markE0TbdrN.resize(n3);
  g.markE0TbdrD.resize(n3);
  g.markV0TbdrD.assign(n3, 0);
  for (size_t i = 0; i < n3; ++i) {
    g.markE0Tint[i] = g.idE0T[i] == 0;
    g.markE0TbdrN[i] = markN ? (markN[i] != 0) : 0;
    g.markE0TbdrD[i] = !(g.markE0Tint[i] || g.markE0TbdrN[i]);
  }
  // markV0TbdrD = ismember(V0T, V0E(E0T(markE0TbdrD), :))
  std::vector<uint8_t> vD(g.numV, 0);
  for (size_t i = 0; i < n3; ++i)
    if (g.markE0TbdrD[i]) {
      int32_t e = g.E0T[i];
      vD[g.V0E[e]] = 1;
      vD[g.V0E[e + g.numE]] = 1;
    }
  for (size_t i = 0; i < n3; ++i) g.markV0TbdrD[i] = vD[g.V0T[i]];
}

void map_ref_to_phy(const Grid& g, int nq, const double* q1, const double* q2, double* X1, double* X2) {
  const int32_t K = g.numT;
  for (int q = 0; q < nq; ++q)
    for (int32_t k = 0; k < K; ++k) {
      double b11 = g.B[k], b21 = g.B[k + K], b12 = g.B[k + 2 * size_t(K)], b22 = g.B[k + 3 * size_t(K)];
      X1[k + size_t(q) * K] = b11 * q1[q] + b12 * q2[q] + g.coordV0T[k];
      X2[k + size_t(q) * K] = b21 * q1[q] + b22 * q2[q] + g.coordV0T[k + 3 * size_t(K)];
    }
}

}  // namespace festung

namespace festung {

namespace {
// Concatenate column-major arrays with `cols` columns; part i has rows[i] rows.
template <class T, class F>
std::vector<T> cat_cols(const std::vector<const Grid*>& gs, const std::vector<int64_t>& rows, int cols,
                        F&& get) {
  int64_t R = 0;
  for (int64_t r : rows) R += r;
  std::vector<T> out(size_t(R) * cols);
  for (int c = 0; c < cols; ++c) {
    int64_t off = 0;
    for (size_t i = 0; i < gs.size(); ++i) {
      for (int64_t r = 0; r < rows[i]; ++r) out[size_t(c) * R + off + r] = get(i, size_t(c) * rows[i] + r);
      off += rows[i];
    }
  }
  return out;
}
}  // namespace

Grid concat_grids(const std::vector<const Grid*>& gs) {
  if (gs.empty()) fail(FESTUNG_ERR_ARG, "concat: no grids");
  const size_t n = gs.size();
  std::vector<int64_t> T(n), V(n), E(n), V1(n);
  std::vector<int32_t> offT(n), offV(n), offE(n);
  int64_t sT = 0, sV = 0, sE = 0;
  for (size_t i = 0; i < n; ++i) {
    T[i] = gs[i]->numT;
    V[i] = gs[i]->numV;
    E[i] = gs[i]->numE;
    offT[i] = int32_t(sT);
    offV[i] = int32_t(sV);
    offE[i] = int32_t(sE);
    sT += T[i];
    sV += V[i];
    sE += E[i];
  }
  if (sT > INT32_MAX / 4 || sV > INT32_MAX / 2) fail(FESTUNG_ERR_ARG, "concat: mesh too large for int32 ids");
  Grid g;
  g.numT = int32_t(sT);
  g.numV = int32_t(sV);
  g.numE = int32_t(sE);
  // part boundaries (parts of parts are kept)
  for (size_t i = 0; i < n; ++i) {
    const Grid& q = *gs[i];
    if (q.partT.size() >= 2) {
      for (size_t j = 0; j + 1 < q.partT.size(); ++j) {
        g.partT.push_back(q.partT[j] + offT[i]);
        g.partV.push_back(q.partV[j] + offV[i]);
      }
    } else {
      g.partT.push_back(offT[i]);
      g.partV.push_back(offV[i]);
    }
  }
  g.partT.push_back(g.numT);
  g.partV.push_back(g.numV);
  auto shift = [](int32_t v, int32_t o) { return v < 0 ? v : v + o; };
  g.coordV = cat_cols<double>(gs, V, 2, [&](size_t i, size_t j) { return gs[i]->coordV[j]; });
  g.V0T = cat_cols<int32_t>(gs, T, 3, [&](size_t i, size_t j) { return shift(gs[i]->V0T[j], offV[i]); });
  g.V0E = cat_cols<int32_t>(gs, E, 2, [&](size_t i, size_t j) { return shift(gs[i]->V0E[j], offV[i]); });
  g.T0E = cat_cols<int32_t>(gs, E, 2, [&](size_t i, size_t j) { return shift(gs[i]->T0E[j], offT[i]); });
  g.E0E = cat_cols<int32_t>(gs, E, 2, [&](size_t i, size_t j) { return gs[i]->E0E[j]; });
  g.E0T = cat_cols<int32_t>(gs, T, 3, [&](size_t i, size_t j) { return shift(gs[i]->E0T[j], offE[i]); });
  g.idE = cat_cols<int32_t>(gs, E, 1, [&](size_t i, size_t j) { return gs[i]->idE[j]; });
  g.idE0T = cat_cols<int32_t>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->idE0T[j]; });
  g.areaE = cat_cols<double>(gs, E, 1, [&](size_t i, size_t j) { return gs[i]->areaE[j]; });
  g.nuE = cat_cols<double>(gs, E, 2, [&](size_t i, size_t j) { return gs[i]->nuE[j]; });
  g.baryE = cat_cols<double>(gs, E, 2, [&](size_t i, size_t j) { return gs[i]->baryE[j]; });
  g.areaT = cat_cols<double>(gs, T, 1, [&](size_t i, size_t j) { return gs[i]->areaT[j]; });
  g.baryT = cat_cols<double>(gs, T, 2, [&](size_t i, size_t j) { return gs[i]->baryT[j]; });
  g.B = cat_cols<double>(gs, T, 4, [&](size_t i, size_t j) { return gs[i]->B[j]; });
  g.areaE0T = cat_cols<double>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->areaE0T[j]; });
  g.sigE0T = cat_cols<double>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->sigE0T[j]; });
  g.nuE0T = cat_cols<double>(gs, T, 6, [&](size_t i, size_t j) { return gs[i]->nuE0T[j]; });
  g.coordV0T = cat_cols<double>(gs, T, 6, [&](size_t i, size_t j) { return gs[i]->coordV0T[j]; });
  g.nbr = cat_cols<int32_t>(gs, T, 3, [&](size_t i, size_t j) { return shift(gs[i]->nbr[j], offT[i]); });
  g.nbrEdge = cat_cols<int32_t>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->nbrEdge[j]; });
  g.markE0Tint = cat_cols<uint8_t>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->markE0Tint[j]; });
  g.markE0TbdrD = cat_cols<uint8_t>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->markE0TbdrD[j]; });
  g.markE0TbdrN = cat_cols<uint8_t>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->markE0TbdrN[j]; });
  g.markV0TbdrD = cat_cols<uint8_t>(gs, T, 3, [&](size_t i, size_t j) { return gs[i]->markV0TbdrD[j]; });
  // vertex patches
  g.v2tPtr.assign(size_t(sV) + 1, 0);
  g.v2tIdx.clear();
  g.v2tIdx.reserve(3 * size_t(sT));
  for (size_t i = 0; i < n; ++i) {
    const Grid& a = *gs[i];
    for (int32_t v = 0; v < a.numV; ++v) {
      for (int32_t t = a.v2tPtr[v]; t < a.v2tPtr[v + 1]; ++t) g.v2tIdx.push_back(a.v2tIdx[t] + offT[i]);
      g.v2tPtr[offV[i] + v + 1] = int32_t(g.v2tIdx.size());
    }
  }
  return g;
}

}  // namespace festung
