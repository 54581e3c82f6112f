// Legacy ASCII VTK output of a displacement field (see vtk.cpp).
#pragma once
#include <string>

#include "problem.hpp"

namespace hfls {

// u: 3 * nPoints global displacements; writes <folder>/box.vtk (iter < 0) or
// <folder>/box<iter>.vtk
void write_vtk(const Problem& P, const double* u, const std::string& folder, int iter);

// MatrixMarket dump of a dense row-major rows x cols matrix A in the layout of
// Matrix::printToFile (Matrix.cpp:556-632): <folder>/dump_<name>_<idx>.txt, header
// "%%MatrixMarket matrix coordinate real general|symmetric", "rows\tcols\tnnz",
// then the nonzero entries "i\tj\tvalue" (1-based, %.16e) row by row; a symmetric
// matrix is written as its upper triangle (i <= j), as the reference stores it.
void write_matrix_market(const std::string& folder, const std::string& name, int idx, int rows, int cols,
                         const double* A, bool symmetric);
// a vector, one value per line (%.16e): <folder>/dump_<name>.txt (Matrix.cpp:531-551)
void write_array(const std::string& folder, const std::string& name, const double* v, int64_t n);

}  // namespace hfls
