// Legacy ASCII VTK output of a displacement field (Mesh::SaveVTK, Mesh.cpp:292-351):
// the points, the hex8 cells (VTK type 12) with the node order of Mesh.cpp:247-263,
// the displacements as a 3-component point scalar, and per cell the partition
// (subdomain id) and material id (Mesh.cpp:271-288: (int)remainder(partition, 2)).
// The section layout and number formats are those of the reference's files so that
// its post-processing reads them unchanged.
#include "vtk.hpp"

#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <vector>

namespace hfls {

void write_vtk(const Problem& P, const double* u, const std::string& folder, int iter) {
  const std::string name = iter < 0 ? folder + "/box.vtk" : folder + "/box" + std::to_string(iter) + ".vtk";
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(name.c_str(), "w"), &std::fclose);
  if (!f) throw std::runtime_error("cannot open " + name + " for writing");
  FILE* o = f.get();
  const int64_t NX = P.NE[0], NY = P.NE[1], NZ = P.NE[2];
  const int64_t nx1 = NX + 1, nxy1 = nx1 * (NY + 1), np = P.nPoints, ne = NX * NY * NZ;
  std::fprintf(o, "# vtk DataFile Version 3.0\nvtk output\nASCII\n\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(o, "POINTS %lld float\n", (long long)np);
  for (int64_t kk = 0; kk <= NZ; kk++)
    for (int64_t jj = 0; jj <= NY; jj++)
      for (int64_t ii = 0; ii <= NX; ii++) {
        double x[3];
        P.node_coords(ii, jj, kk, x);
        std::fprintf(o, "%f %f %f\n", x[0], x[1], x[2]);
      }
  std::fprintf(o, "CELLS %lld %lld\n", (long long)ne, (long long)(9 * ne));
  for (int64_t kk = 0; kk < NZ; kk++)
    for (int64_t jj = 0; jj < NY; jj++)
      for (int64_t ii = 0; ii < NX; ii++) {
        const int64_t a = ii + jj * nx1 + kk * nxy1;
        const int64_t q[4] = {a, a + 1, a + 1 + nx1, a + nx1};
        std::fprintf(o, "%d ", 8);
        for (int l = 0; l < 8; l++) std::fprintf(o, "% lld", (long long)(q[l & 3] + (l >> 2) * nxy1));
        std::fprintf(o, "\n");
      }
  std::fprintf(o, "CELL_TYPES %lld\n", (long long)ne);
  for (int64_t e = 0; e < ne; e++) std::fprintf(o, "%d\n", 12);
  std::fprintf(o, "POINT_DATA %lld\nSCALARS displacements float 3\nLOOKUP_TABLE my_table\n", (long long)np);
  for (int64_t p = 0; p < np; p++) std::fprintf(o, "%f %f %f\n", u[3 * p], u[3 * p + 1], u[3 * p + 2]);
  // partition of every element from its slab indices
  std::vector<int> slab[3];
  for (int a = 0; a < 3; a++) {
    slab[a].resize(P.NE[a]);
    for (int s = 0; s + 1 < (int)P.starts[a].size(); s++)
      for (int e = P.starts[a][s]; e < P.starts[a][s + 1]; e++) slab[a][e] = s;
  }
  std::vector<int> part((size_t)ne);
  for (int64_t kk = 0, e = 0; kk < NZ; kk++)
    for (int64_t jj = 0; jj < NY; jj++)
      for (int64_t ii = 0; ii < NX; ii++, e++) part[e] = P.sub_id(slab[0][ii], slab[1][jj], slab[2][kk]);
  std::fprintf(o, "\nCELL_DATA %lld\nSCALARS decomposition int 1\nLOOKUP_TABLE decomposition\n", (long long)ne);
  for (int64_t e = 0; e < ne; e++) std::fprintf(o, "%d\n", part[e]);
  // material id from the partition createMesh made (the geometric one also after
  // metis=true merged it, Mesh.cpp:271-288)
  std::fprintf(o, "SCALARS materialId int 1\nLOOKUP_TABLE materialId\n");
  for (int64_t kk = 0; kk < NZ; kk++)
    for (int64_t jj = 0; jj < NY; jj++)
      for (int64_t ii = 0; ii < NX; ii++)
        std::fprintf(o, "%d\n", (int)std::remainder((double)P.material_partition(ii, jj, kk), 2.0));
  if (std::ferror(o)) throw std::runtime_error("write error on " + name);
}

void write_matrix_market(const std::string& folder, const std::string& name, int idx, int rows, int cols,
                         const double* A, bool symmetric) {
  const std::string path = folder + "/dump_" + name + "_" + std::to_string(idx) + ".txt";
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path.c_str(), "w"), &std::fclose);
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  int64_t nnz = 0;
  for (int i = 0; i < rows; i++)
    for (int j = symmetric ? i : 0; j < cols; j++)
      if (A[(size_t)i * cols + j] != 0.0) nnz++;
  std::fprintf(f.get(), "%%%%MatrixMarket matrix coordinate real %s\n", symmetric ? "symmetric" : "general");
  std::fprintf(f.get(), "%d\t%d\t%lld\n", rows, cols, (long long)nnz);
  for (int i = 0; i < rows; i++)
    for (int j = symmetric ? i : 0; j < cols; j++) {
      const double v = A[(size_t)i * cols + j];
      if (v != 0.0) std::fprintf(f.get(), "%d\t%d\t%.16e\n", i + 1, j + 1, v);
    }
}

void write_array(const std::string& folder, const std::string& name, const double* v, int64_t n) {
  const std::string path = folder + "/dump_" + name + ".txt";
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path.c_str(), "w"), &std::fclose);
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  for (int64_t i = 0; i < n; i++) std::fprintf(f.get(), "%.16e\n", v[i]);
}

}  // namespace hfls
