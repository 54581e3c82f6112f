// Host-side problem description: structured cube mesh, subdomain boxes, the
// fully redundant gluing Bf with its weights, and the per-subdomain orderings the
// device factorisation uses.  All of this is integer work that must agree bit for
// bit with the reference's construction:
//   mesh numbering        Mesh.cpp:152-171, 243-288
//   corner / Dirichlet    Mesh.cpp:175-238
//   l2g (sorted DOFs)     Data.cpp:111-144
//   Bf rows and weights   Cluster.cpp:846-924 (order d, then DOF, then neighbour)
//   compressed rows       Cluster.cpp:931-959, j_col_cmpr Cluster.cpp:756-765
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace hfls {

using Options = std::map<std::string, std::string>;

Options default_options();
std::string opt(const Options& o, const std::string& k);
double optd(const Options& o, const std::string& k);
int opti(const Options& o, const std::string& k);

struct Subdomain {
  int id = 0, cluster = 0;
  int grid[3] = {0, 0, 0};     // global slab index per axis
  int e0[3] = {0, 0, 0};       // first global element per axis
  int ne[3] = {0, 0, 0};       // elements per axis
  int nn[3] = {0, 0, 0};       // local nodes per axis
  int n = 0;                   // DOFs
  double E = 0.0;
  // compressed Bf (rows ascending): global row, local DOF, sign
  std::vector<int> bf_row, bf_dof;
  std::vector<int8_t> bf_sign;
  std::vector<int> bdofs;      // j_col_cmpr: sorted B-touched local DOFs
  std::vector<int> sset;       // kernel-detection s-set (sorted local DOFs, subset of bdofs)
  // device orderings
  std::vector<int> rdofs;      // interior (not B-touched), ascending
  std::vector<int> border;     // b-order: [b \ s ascending, s ascending]
  std::vector<int> bpos;       // local DOF -> position in b-order, -1 for interior
  int n_r = 0, n_s = 0, n_bs = 0, n_ss = 0;
  std::vector<uint8_t> owner;  // per local DOF: 1 if this subdomain writes u there
  int64_t global_dof(int ldof) const;
};

struct Problem {
  Options o;
  int C[3], N[3], NE[3];
  std::vector<int> slabs[3], starts[3];
  int nsub = 0, nsub_cl = 0, ncl = 0;
  int64_t nPoints = 0;
  double dx[3];
  std::vector<Subdomain> subs;
  std::vector<int64_t> corner_dofs, dirichlet_dofs;
  std::vector<uint8_t> is_dir;   // per global DOF
  // lambda rows: entry a (sign +1) and b (sign -1, sub -1 for Dirichlet rows)
  int64_t nlam = 0, n_glue = 0;
  std::vector<int> row_sub_a, row_sub_b;
  std::vector<int64_t> row_dof;   // global DOF of the row
  std::vector<double> weight;
  // metis=true: the geometric Nx x Ny x Nz partition that createMesh made before
  // ddm_metis merged it (geo_n elements per partition and axis, geo_N partitions
  // along x and y; zeros otherwise).  It keeps the material ids (Mesh.cpp:271-288):
  // with the reference material an element of an odd geometric partition has
  // mat_E1, an even one its subdomain's E (mat_per_element)
  int geo_n[3] = {0, 0, 0}, geo_N[2] = {0, 0};
  bool mat_per_element = false;
  double mat_E1 = 0.0;
  // geometric partition of element (ii, jj, kk): geo_n set, else the subdomain id
  int material_partition(int64_t ii, int64_t jj, int64_t kk) const;

  explicit Problem(const Options& o);
  int sub_id(int I, int J, int K) const;
  // subdomains containing global node (ii,jj,kk), ascending
  int subs_of_node(int64_t ii, int64_t jj, int64_t kk, int* out) const;
  void node_coords(int64_t ii, int64_t jj, int64_t kk, double* x) const;
  double young(int d) const;
};

// splitmix64 stream shared with the oracle (material, loads)
uint64_t splitmix64(uint64_t& state);
double uniform01(uint64_t& state);

// k seeded random smoothed tractions on the x = 1 face as packed local loads
// (column j at j * sum_d n_d) for load cases case0 .. case0 + k - 1
void traction_loads(const Problem& P, uint64_t seed, int case0, int k, double* f);
// the smoothed traction field t[(c * (NZ + 1) + kk) * (NY + 1) + jj] of load case lcase
// on the x = 1 face (the first part of traction_loads; the lumping to subdomain
// loads can then run on the device)
void traction_face(const Problem& P, uint64_t seed, int lcase, double* t);

}  // namespace hfls
