// Unstructured triangle mesh with the reference's numbering and derived
// connectivity (host side).
#pragma once
#include <cstdint>
#include <vector>

namespace festung {

// All arrays are column-major with the reference's shapes (e.g. E0T is
// numT x 3 stored as E0T[k + n*numT]); ids are 0-based, "none" is -1.
struct Grid {
  int32_t numT = 0, numV = 0, numE = 0, dim = 0;
  std::vector<double> coordV;              // numV x 2
  std::vector<int32_t> V0T;                // numT x 3
  std::vector<int32_t> V0E, T0E, E0E;      // numE x 2
  std::vector<int32_t> E0T;                // numT x 3
  std::vector<int32_t> idE;                // numE (boundary ids 1..4, 0 = none)
  std::vector<int32_t> idE0T;              // numT x 3
  std::vector<double> areaE, nuE, baryE;   // numE, numE x 2, numE x 2
  std::vector<double> areaT, baryT;        // numT, numT x 2
  std::vector<double> B;                   // numT x 2 x 2, B(k, m, j)
  std::vector<double> areaE0T;             // numT x 3
  std::vector<double> nuE0T;               // numT x 3 x 2
  std::vector<double> sigE0T;              // numT x 3
  std::vector<double> coordV0T;            // numT x 3 x 2
  std::vector<int32_t> nbr, nbrEdge;       // numT x 3: neighbour element / its local edge (markE0TE0T)
  std::vector<int32_t> v2tPtr, v2tIdx;     // vertex -> element CSR (vertex patches)
  // boundary marks (advection/configureProblem.m:148-150,
  // advection/preprocessProblem.m:72-76)
  std::vector<uint8_t> markE0Tint, markE0TbdrD, markE0TbdrN, markV0TbdrD;  // numT x 3
  // concat_grids: first element / vertex of each part, then the totals
  // (empty for a grid of one part)
  std::vector<int32_t> partT, partV;
};

// core/generateGridData.m:130-198; V0T 0-based, numT x 3 column-major
Grid generate_grid_data(const double* coordV, int32_t numV, const int32_t* V0T, int32_t numT);
// core/domainSquare.m:67-105
Grid domain_square(double h, double a, double b);
// Boundary marks for given boundary ids and Neumann marks.
void mark_boundaries(Grid& g, const uint8_t* markE0TbdrN /* may be null */);
// Disjoint union of several grids (ids offset in order); used to run many
// independent problems of the same degree as one device problem.
Grid concat_grids(const std::vector<const Grid*>& gs);
// Physical coordinates of reference points: X{1,2} numT x nq (mapRef2Phy).
void map_ref_to_phy(const Grid& g, int nq, const double* q1, const double* q2, double* X1, double* X2);

}  // namespace festung

// Opaque handle of the C-ABI.
struct festung_grid {
  festung::Grid g;
};
