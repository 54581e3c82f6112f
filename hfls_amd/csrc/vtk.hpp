// Legacy ASCII VTK output of a displacement field (see vtk.cpp).
#pragma once
#include <string>

#include "problem.hpp"

namespace hfls {

// u: 3 * nPoints global displacements; writes <folder>/box.vtk (iter < 0) or
// <folder>/box<iter>.vtk
void write_vtk(const Problem& P, const double* u, const std::string& folder, int iter);

}  // namespace hfls
