// Host-side integer work of the HTFETI path (see problem.hpp for the reference
// lines each part follows).  Built with -ffp-contract=off so that coordinate
// arithmetic rounds exactly as the reference's (x = (ii*dx - 0.5) + 0.5).
#include "problem.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <sstream>
#include <stdexcept>

namespace hfls {

Options default_options() {
  return Options{
      {"Nx", "2"}, {"Ny", "2"}, {"Nz", "2"}, {"nx", "5"}, {"ny", "5"}, {"nz", "5"},
      {"Cx", "1"}, {"Cy", "1"}, {"Cz", "1"},
      {"young_modulus", "1000"}, {"ratio_mat", "1e4"}, {"poissons_ratio", "0.3"},
      {"material", "reference"}, {"mat_sigma", "1.5"}, {"mat_contrast", "1e3"}, {"seed", "0"},
      {"bc", "symmetry"}, {"load", "reference"}, {"body_force", "-1.0"}, {"density", "2.0"},
      {"typeBc", "ker"}, {"Bc_fullRank", "true"}, {"preconditioner", "Dirichlet_explicit"},
      {"eps_iter", "1e-6"}, {"max_iter", "200"}, {"Ac_extended_by_kerGc", "true"},
      {"slabs_x", ""}, {"slabs_y", ""}, {"slabs_z", ""},
      {"linear_solver", "hip"}, {"vtkWithinIter", "false"}, {"print_matrices", "0"},
      {"precision", "fp64"}, {"nrhs", "1"}, {"device", "0"}, {"refine_tol", "1e-4"}, {"slab_choices", ""},
      {"block_gemm", "f16x2"},
      // the rest of the reference's options map (htfeti.cpp:12-57)
      {"metis", "false"}, {"nparts", "1"}, {"path2data", ""}, {"Ac_clust_explicit", "true"},
      {"GcTGc_assembl_block_by_block", "true"}, {"create_analytic_ker_K", "false"}};
}

std::string opt(const Options& o, const std::string& k) {
  auto it = o.find(k);
  if (it != o.end()) return it->second;
  Options d = default_options();
  auto jt = d.find(k);
  if (jt == d.end()) throw std::runtime_error("unknown option " + k);
  return jt->second;
}
double optd(const Options& o, const std::string& k) { return std::strtod(opt(o, k).c_str(), nullptr); }
int opti(const Options& o, const std::string& k) { return std::atoi(opt(o, k).c_str()); }

uint64_t splitmix64(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ULL;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) {
  uint64_t z = splitmix64(s);
  return ((double)(z >> 11) + 0.5) / 9007199254740992.0;
}

static std::vector<int> parse_slabs(const Options& o, char axis) {
  std::string a(1, axis);
  int C = opti(o, std::string("C") + a), N = opti(o, std::string("N") + a);
  std::string s = opt(o, "slabs_" + a);
  std::vector<int> v;
  const std::string choices = opt(o, "slab_choices");
  if (!s.empty()) {
    std::stringstream ss(s);
    std::string t;
    while (std::getline(ss, t, ',')) v.push_back(std::atoi(t.c_str()));
    if ((int)v.size() != C * N) throw std::runtime_error("slabs_" + a + " needs C*N entries");
  } else if (!choices.empty()) {
    // seeded draw of every slab's element count from the list (uneven batching):
    // slab i of axis a takes choices[floor(u * #choices)], u from a splitmix64
    // stream seeded with seed ^ (0xC2B2AE3D27D4EB4F * (a + 1)), a = 0, 1, 2 for x, y, z
    std::vector<int> ch;
    std::stringstream ss(choices);
    std::string t;
    while (std::getline(ss, t, ',')) ch.push_back(std::atoi(t.c_str()));
    if (ch.empty()) throw std::runtime_error("slab_choices is empty");
    const uint64_t ax = (uint64_t)(axis - 'x') + 1;
    uint64_t st = (uint64_t)std::strtoull(opt(o, "seed").c_str(), nullptr, 10) ^ (0xC2B2AE3D27D4EB4FULL * ax);
    for (int i = 0; i < C * N; i++) {
      const int idx = (int)(uniform01(st) * (double)ch.size());
      v.push_back(ch[std::min(idx, (int)ch.size() - 1)]);
    }
  } else {
    v.assign(C * N, opti(o, std::string("n") + a));
  }
  for (int x : v)
    if (x < 1) throw std::runtime_error("slab element count must be >= 1");
  return v;
}

int64_t Subdomain::global_dof(int ldof) const { return -1; }

int Problem::sub_id(int I, int J, int K) const {
  int cl = (I / N[0]) + (J / N[1]) * C[0] + (K / N[2]) * C[0] * C[1];
  int loc = (I % N[0]) + (J % N[1]) * N[0] + (K % N[2]) * N[0] * N[1];
  return cl * nsub_cl + loc;
}

// slabs along one axis containing grid index i (1 or 2, ascending)
static int slabs_of(const std::vector<int>& st, int64_t i, int* out) {
  int ns = (int)st.size() - 1;
  int I = (int)(std::upper_bound(st.begin(), st.end(), i) - st.begin()) - 1;
  if (I >= ns) I = ns - 1;
  int c = 0;
  if (I > 0 && st[I] == i) out[c++] = I - 1;
  out[c++] = I;
  return c;
}

int Problem::subs_of_node(int64_t ii, int64_t jj, int64_t kk, int* out) const {
  int sx[2], sy[2], sz[2];
  int cx = slabs_of(starts[0], ii, sx), cy = slabs_of(starts[1], jj, sy), cz = slabs_of(starts[2], kk, sz);
  int c = 0;
  for (int a = 0; a < cz; a++)
    for (int b = 0; b < cy; b++)
      for (int e = 0; e < cx; e++) out[c++] = sub_id(sx[e], sy[b], sz[a]);
  std::sort(out, out + c);
  return c;
}

void Problem::node_coords(int64_t ii, int64_t jj, int64_t kk, double* x) const {
  // Mesh.cpp:158-168 with shift (0.5, 0.5, 0) and flag_radius = false
  volatile double t0 = (double)ii * dx[0] - 0.5;
  volatile double t1 = (double)jj * dx[1] - 0.5;
  volatile double t2 = (double)kk * dx[2];
  x[0] = t0 + 0.5;
  x[1] = t1 + 0.5;
  x[2] = t2 + 0.0;
}

double Problem::young(int d) const { return subs[d].E; }

int Problem::material_partition(int64_t ii, int64_t jj, int64_t kk) const {
  if (geo_n[0] > 0)
    return (int)(ii / geo_n[0] + (jj / geo_n[1]) * geo_N[0] + (kk / geo_n[2]) * geo_N[0] * geo_N[1]);
  const int I = (int)(std::upper_bound(starts[0].begin(), starts[0].end(), ii) - starts[0].begin()) - 1;
  const int J = (int)(std::upper_bound(starts[1].begin(), starts[1].end(), jj) - starts[1].begin()) - 1;
  const int K = (int)(std::upper_bound(starts[2].begin(), starts[2].end(), kk) - starts[2].begin()) - 1;
  return sub_id(I, J, K);
}

static bool is_bool(const std::string& v) { return v == "true" || v == "false"; }

// The keys of the reference's options map that select code paths (htfeti.cpp:12-57):
// every value the reference accepts is either honoured or rejected with the reason.
static void check_reference_options(const Options& o) {
  // linear_solver picks the sparse direct solver behind Matrix::sym_factor /
  // num_factor / solve_system (Matrix.cpp:970-985, 1205-1233); here that seam is
  // the device Cholesky for both the reference's default (pardiso) and hip
  const std::string ls = opt(o, "linear_solver");
  if (ls == "dissection")
    throw std::runtime_error("linear_solver=dissection (Matrix.cpp:886-968) is not available; use pardiso or hip");
  if (ls != "pardiso" && ls != "hip") throw std::runtime_error("unknown linear_solver " + ls + " (pardiso or hip)");
  const std::string me = opt(o, "metis");
  if (!is_bool(me)) throw std::runtime_error("metis must be true or false");
  if (me == "true" && opti(o, "nparts") > 1)
    throw std::runtime_error(
        "metis=true with nparts > 1 partitions the mesh with METIS_PartMeshDual (Mesh.cpp:392-412), which this "
        "library does not provide; use metis=false (geometric Nx x Ny x Nz decomposition) or nparts=1");
  // Ac_clust_explicit=false: the Sc branch (Cluster.cpp:325-342, Sc = Gc^T Fc^-1 Gc),
  // completed as the two-stage coarse solve (myModul.py:128-135)
  if (!is_bool(opt(o, "Ac_clust_explicit"))) throw std::runtime_error("Ac_clust_explicit must be true or false");
  // GcTGc_assembl_block_by_block picks one of two assemblies of the same Gc^T Gc
  // (Cluster.cpp:236-245); create_analytic_ker_K is read by no solver branch
  // (Cluster.cpp:84-86 is `if (0)`), and htfeti.cpp:60-66 can leave it empty
  for (const char* k : {"Ac_extended_by_kerGc", "GcTGc_assembl_block_by_block", "Bc_fullRank", "vtkWithinIter"})
    if (!is_bool(opt(o, k))) throw std::runtime_error(std::string(k) + " must be true or false");
  // print_matrices 1: MatrixMarket dumps of the cluster matrices (Cluster.cpp:220-277);
  // 2 and 3 add per-subdomain CSR dumps of K, Bc, Bf, R, Gc, Rf, Gf (Cluster.cpp:126-212,
  // 362-370), which this library does not write (it never forms K_i in CSR)
  const int pm = opti(o, "print_matrices");
  if (pm < 0 || pm > 3) throw std::runtime_error("print_matrices must be 0..3");
  if (pm > 1)
    throw std::runtime_error("print_matrices > 1 (per-subdomain CSR dumps, Cluster.cpp:126-212) is not supported; "
                             "use 0 or 1");
  if (pm > 0 && opt(o, "path2data").empty()) throw std::runtime_error("print_matrices > 0 needs path2data");
}

Problem::Problem(const Options& oin) : o(oin) {
  // htfeti.cpp:18 documents Dirichlet_explicit, Dirichlet_implicit and lumped.
  // Both Dirichlet forms apply S_bb (Cluster.cpp:1061-1102; explicit and implicit
  // are equal in exact arithmetic, here S_bb is always explicit).  The reference
  // takes any other string as lumped (Cluster.cpp:1104-1107); a misspelt value is
  // rejected here instead of silently switching the preconditioner.
  {
    const std::string prec = opt(o, "preconditioner");
    if (prec != "Dirichlet_explicit" && prec != "Dirichlet_implicit" && prec != "lumped")
      throw std::runtime_error("unknown preconditioner " + prec +
                               " (Dirichlet_explicit, Dirichlet_implicit or lumped)");
    const std::string tb = opt(o, "typeBc");
    if (tb != "ker" && tb != "cor" && tb != "all" && tb != "cor+edge")
      throw std::runtime_error("unknown typeBc " + tb + " (ker, cor, all or cor+edge)");
    check_reference_options(o);
    // matrix-core format of K+_bb and S_i in mixed-precision blocks: split FP16
    // (hi + lo planes) or FP32
    const std::string bg = opt(o, "block_gemm");
    if (bg != "f16x2" && bg != "f32") throw std::runtime_error("unknown block_gemm " + bg + " (f16x2 or f32)");
  }
  const char* ax = "xyz";
  for (int a = 0; a < 3; a++) {
    C[a] = opti(o, std::string("C") + ax[a]);
    N[a] = opti(o, std::string("N") + ax[a]);
    slabs[a] = parse_slabs(o, ax[a]);
    starts[a].assign(1, 0);
    for (int v : slabs[a]) starts[a].push_back(starts[a].back() + v);
    NE[a] = starts[a].back();
    dx[a] = 1.0 / NE[a];
  }
  const int64_t nx1 = NE[0] + 1, ny1 = NE[1] + 1, nz1 = NE[2] + 1;
  nPoints = nx1 * ny1 * nz1;
  const int64_t nxy1 = nx1 * ny1;

  // corner DOFs (Mesh.cpp:175-208): from the geometric Nx x Ny x Nz grid, which
  // createMesh uses before any METIS partition
  {
    std::set<int64_t> cube = {0, NE[0], NE[1] * nx1, NE[1] * nx1 + NE[0]};
    std::set<int64_t> top;
    for (int64_t c : cube) top.insert(c + NE[2] * nxy1);
    cube.insert(top.begin(), top.end());
    for (int k : starts[2])
      for (int j : starts[1])
        for (int i : starts[0]) {
          int64_t n = i + nx1 * j + nxy1 * k;
          if (cube.count(n)) continue;
          for (int c = 0; c < 3; c++) corner_dofs.push_back(3 * n + c);
        }
  }
  // metis=true with nparts <= 1: Mesh::ddm_metis puts every element in partition 0
  // without calling METIS (Mesh.cpp:413-421), i.e. the whole cube is one subdomain
  if (opt(o, "metis") == "true") {
    if (C[0] * C[1] * C[2] != 1) throw std::runtime_error("metis=true needs one cluster (Cx = Cy = Cz = 1)");
    if (!opt(o, "slabs_x").empty() || !opt(o, "slabs_y").empty() || !opt(o, "slabs_z").empty() ||
        !opt(o, "slab_choices").empty())
      throw std::runtime_error("metis=true takes the uniform Nx x Ny x Nz grid of nx x ny x nz elements");
    // the material ids stay those of the geometric partition (createMesh sets them,
    // ddm_metis changes only the partition ids, Mesh.cpp:271-288, 425-427)
    for (int a = 0; a < 3; a++) geo_n[a] = slabs[a][0];
    geo_N[0] = N[0];
    geo_N[1] = N[1];
    if (opt(o, "material") == "reference") {
      mat_per_element = true;
      mat_E1 = optd(o, "young_modulus");
    }
    for (int a = 0; a < 3; a++) {
      N[a] = 1;
      slabs[a].assign(1, NE[a]);
      starts[a] = {0, NE[a]};
    }
  }
  nsub_cl = N[0] * N[1] * N[2];
  ncl = C[0] * C[1] * C[2];
  nsub = nsub_cl * ncl;
  // Dirichlet DOFs (Mesh.cpp:221-238), points ascending, components x,y,z
  {
    std::string bc = opt(o, "bc");
    if (bc != "symmetry" && bc != "clamp_x0") throw std::runtime_error("unknown bc " + bc);
    is_dir.assign(3 * nPoints, 0);
    const double eps0 = 1e-8;
    for (int64_t kk = 0; kk < nz1; kk++)
      for (int64_t jj = 0; jj < ny1; jj++)
        for (int64_t ii = 0; ii < nx1; ii++) {
          double x[3];
          node_coords(ii, jj, kk, x);
          int64_t p = ii + jj * nx1 + kk * nxy1;
          if (bc == "symmetry") {
            for (int c = 0; c < 3; c++)
              if (std::fabs(x[c]) < eps0) dirichlet_dofs.push_back(3 * p + c);
          } else if (std::fabs(x[0]) < eps0) {
            for (int c = 0; c < 3; c++) dirichlet_dofs.push_back(3 * p + c);
          }
        }
    for (int64_t g : dirichlet_dofs) is_dir[g] = 1;
  }
  // materials (Data.cpp:36,47-48; Mesh.cpp:281 for the reference checkerboard)
  std::vector<double> Es(nsub);
  {
    std::string mode = opt(o, "material");
    double E = optd(o, "young_modulus");
    if (mode == "reference") {
      double r = optd(o, "ratio_mat");
      for (int d = 0; d < nsub; d++) Es[d] = ((int)std::remainder((double)d, 2.0) == 0) ? E * r : E;
    } else if (mode == "uniform") {
      for (int d = 0; d < nsub; d++) Es[d] = E;
    } else if (mode == "lognormal") {
      double sig = optd(o, "mat_sigma"), half = std::log(optd(o, "mat_contrast")) / 2.0;
      uint64_t st = (uint64_t)std::strtoull(opt(o, "seed").c_str(), nullptr, 10);
      for (int d = 0; d < nsub; d++) {
        double u0 = uniform01(st), u1 = uniform01(st);
        double z = std::sqrt(-2.0 * std::log(u0)) * std::cos(2.0 * M_PI * u1);
        double t = std::min(std::max(sig * z, -half), half);
        Es[d] = E * std::exp(t);
      }
    } else {
      throw std::runtime_error("unknown material " + mode);
    }
  }
  // subdomain boxes
  subs.resize(nsub);
  for (int K = 0; K < (int)slabs[2].size(); K++)
    for (int J = 0; J < (int)slabs[1].size(); J++)
      for (int I = 0; I < (int)slabs[0].size(); I++) {
        int d = sub_id(I, J, K);
        Subdomain& s = subs[d];
        s.id = d;
        s.cluster = d / nsub_cl;
        int g3[3] = {I, J, K};
        for (int a = 0; a < 3; a++) {
          s.grid[a] = g3[a];
          s.e0[a] = starts[a][g3[a]];
          s.ne[a] = slabs[a][g3[a]];
          s.nn[a] = s.ne[a] + 1;
        }
        s.n = 3 * s.nn[0] * s.nn[1] * s.nn[2];
        s.E = Es[d];
      }
  // Bf glue rows (Cluster.cpp:846-903) then Dirichlet rows (Cluster.cpp:906-924)
  int list[8];
  for (int d = 0; d < nsub; d++) {
    Subdomain& s = subs[d];
    for (int ld = 0; ld < s.n; ld++) {
      int ln = ld / 3, c = ld % 3;
      int64_t ii = s.e0[0] + ln % s.nn[0], jj = s.e0[1] + (ln / s.nn[0]) % s.nn[1],
              kk = s.e0[2] + ln / (s.nn[0] * s.nn[1]);
      int64_t g = 3 * (ii + jj * nx1 + kk * nxy1) + c;
      if (is_dir[g]) continue;
      int m = subs_of_node(ii, jj, kk, list);
      if (m < 2) continue;
      for (int q = 0; q < m; q++) {
        int t = list[q];
        if (t <= d) continue;
        Subdomain& u = subs[t];
        int lnu = (int)((ii - u.e0[0]) + (jj - u.e0[1]) * u.nn[0] + (kk - u.e0[2]) * u.nn[0] * u.nn[1]);
        int64_t r = (int64_t)row_dof.size();
        row_dof.push_back(g);
        row_sub_a.push_back(d);
        row_sub_b.push_back(t);
        weight.push_back(1.0 / ((double)(m - 1) + 1.0));
        s.bf_row.push_back((int)r); s.bf_dof.push_back(ld); s.bf_sign.push_back(1);
        u.bf_row.push_back((int)r); u.bf_dof.push_back(3 * lnu + c); u.bf_sign.push_back(-1);
      }
    }
  }
  n_glue = (int64_t)row_dof.size();
  for (int d = 0; d < nsub; d++) {
    Subdomain& s = subs[d];
    for (int ld = 0; ld < s.n; ld++) {
      int ln = ld / 3, c = ld % 3;
      int64_t ii = s.e0[0] + ln % s.nn[0], jj = s.e0[1] + (ln / s.nn[0]) % s.nn[1],
              kk = s.e0[2] + ln / (s.nn[0] * s.nn[1]);
      int64_t g = 3 * (ii + jj * nx1 + kk * nxy1) + c;
      if (!is_dir[g]) continue;
      int64_t r = (int64_t)row_dof.size();
      row_dof.push_back(g);
      row_sub_a.push_back(d);
      row_sub_b.push_back(-1);
      weight.push_back(1.0);
      s.bf_row.push_back((int)r); s.bf_dof.push_back(ld); s.bf_sign.push_back(1);
    }
  }
  nlam = (int64_t)row_dof.size();
  if (nlam >= (int64_t)1 << 31) throw std::runtime_error("too many lambda rows");

  // per-subdomain B-touched set, s-set, orderings, owners
  for (int d = 0; d < nsub; d++) {
    Subdomain& s = subs[d];
    std::vector<uint8_t> tb(s.n, 0);
    for (int ld : s.bf_dof) tb[ld] = 1;
    for (int ld = 0; ld < s.n; ld++)
      if (tb[ld]) s.bdofs.push_back(ld);
    if (s.bdofs.empty()) throw std::runtime_error("subdomain without gluing or Dirichlet rows");
    // s-set: extreme B-touched nodes along the 8 diagonals (integer grid coords)
    std::vector<int> bnodes;
    for (int ld : s.bdofs)
      if (bnodes.empty() || bnodes.back() != ld / 3) bnodes.push_back(ld / 3);
    std::set<int> chosen;
    for (int sz = 1; sz >= -1; sz -= 2)
      for (int sy = 1; sy >= -1; sy -= 2)
        for (int sx = 1; sx >= -1; sx -= 2) {
          int64_t best = 0;
          int bn = -1;
          for (int ln : bnodes) {
            int64_t ii = s.e0[0] + ln % s.nn[0], jj = s.e0[1] + (ln / s.nn[0]) % s.nn[1],
                    kk = s.e0[2] + ln / (s.nn[0] * s.nn[1]);
            int64_t v = sx * ii + sy * jj + sz * kk;
            if (bn < 0 || v > best) { best = v; bn = ln; }
          }
          chosen.insert(bn);
        }
    for (int ln : chosen)
      for (int c = 0; c < 3; c++)
        if (tb[3 * ln + c]) s.sset.push_back(3 * ln + c);
    std::sort(s.sset.begin(), s.sset.end());
    if ((int)s.sset.size() < 7 || (int)s.sset.size() > 32)
      throw std::runtime_error("kernel-detection s-set has an unsupported size");
    std::vector<uint8_t> ins(s.n, 0);
    for (int ld : s.sset) ins[ld] = 1;
    s.bpos.assign(s.n, -1);
    for (int ld = 0; ld < s.n; ld++)
      if (!tb[ld]) s.rdofs.push_back(ld);
    for (int ld : s.bdofs)
      if (!ins[ld]) s.border.push_back(ld);
    for (int ld : s.sset) s.border.push_back(ld);
    for (int p = 0; p < (int)s.border.size(); p++) s.bpos[s.border[p]] = p;
    s.n_r = (int)s.rdofs.size();
    s.n_s = (int)s.border.size();
    s.n_ss = (int)s.sset.size();
    s.n_bs = s.n_s - s.n_ss;
    if (s.n_r < 1) throw std::runtime_error("subdomain without interior DOFs");
    s.owner.assign(s.n, 0);
    for (int ln = 0; ln < s.n / 3; ln++) {
      int64_t ii = s.e0[0] + ln % s.nn[0], jj = s.e0[1] + (ln / s.nn[0]) % s.nn[1],
              kk = s.e0[2] + ln / (s.nn[0] * s.nn[1]);
      int m = subs_of_node(ii, jj, kk, list);
      if (list[m - 1] == d)
        for (int c = 0; c < 3; c++) s.owner[3 * ln + c] = 1;
    }
  }
}

// Random smoothed surface traction on the x = 1 face (no counterpart in the
// reference, whose only load is the body force of Data.cpp:315-318):
//   * t(c, jj, kk) ~ N(0,1) per face node and component, Box-Muller over a
//     splitmix64 stream seeded with seed ^ (0xD1B54A32D192ED03 * (case + 1)),
//     drawn kk outer, jj, then component (two uniforms per normal);
//   * three Jacobi passes of the 5-point face smoother
//       t(p) <- (t(p) + t(jj-1) + t(jj+1) + t(kk-1) + t(kk+1)) / (1 + #neighbours);
//   * lumped nodal forces per subdomain: f = t * (cnt * hy * hz * 0.25), cnt the
//     number of the subdomain's face elements at the node, hy = 1/NY, hz = 1/NZ.
// Output: k packed local vectors (column j at j * sum_d n_d), cases case0..case0+k-1.
void traction_face(const Problem& P, uint64_t seed, int lcase, double* t) {
  const int NY = P.NE[1], NZ = P.NE[2];
  const int ny1 = NY + 1, nz1 = NZ + 1;
  std::vector<double> u((size_t)3 * ny1 * nz1), cnt((size_t)ny1 * nz1);
  auto at = [&](int c, int jj, int kk) -> size_t { return ((size_t)c * nz1 + kk) * ny1 + jj; };
  const uint64_t cs = (uint64_t)lcase + 1;
  uint64_t st = seed ^ (0xD1B54A32D192ED03ULL * cs);
  for (int kk = 0; kk < nz1; kk++)
    for (int jj = 0; jj < ny1; jj++)
      for (int c = 0; c < 3; c++) {
        double u0 = uniform01(st), u1 = uniform01(st);
        t[at(c, jj, kk)] = std::sqrt(-2.0 * std::log(u0)) * std::cos(2.0 * M_PI * u1);
      }
  for (int kk = 0; kk < nz1; kk++)
    for (int jj = 0; jj < ny1; jj++) {
      double m = 1.0;
      if (jj > 0) m += 1.0;
      if (jj < NY) m += 1.0;
      if (kk > 0) m += 1.0;
      if (kk < NZ) m += 1.0;
      cnt[(size_t)kk * ny1 + jj] = m;
    }
  double* a = t;
  double* b = u.data();
  for (int pass = 0; pass < 3; pass++) {
    for (int c = 0; c < 3; c++)
      for (int kk = 0; kk < nz1; kk++)
        for (int jj = 0; jj < ny1; jj++) {
          double v = a[at(c, jj, kk)];
          if (jj > 0) v += a[at(c, jj - 1, kk)];
          if (jj < NY) v += a[at(c, jj + 1, kk)];
          if (kk > 0) v += a[at(c, jj, kk - 1)];
          if (kk < NZ) v += a[at(c, jj, kk + 1)];
          b[at(c, jj, kk)] = v / cnt[(size_t)kk * ny1 + jj];
        }
    std::swap(a, b);
  }
  if (a != t) std::copy(a, a + u.size(), t);  // three passes: the result is in u
}

void traction_loads(const Problem& P, uint64_t seed, int case0, int k, double* f) {
  const int NY = P.NE[1], NZ = P.NE[2], NX = P.NE[0];
  const int ny1 = NY + 1, nz1 = NZ + 1;
  const double hy = 1.0 / NY, hz = 1.0 / NZ;
  int64_t np = 0;
  for (const Subdomain& s : P.subs) np += s.n;
  std::vector<double> t((size_t)3 * ny1 * nz1);
  auto at = [&](int c, int jj, int kk) -> size_t { return ((size_t)c * nz1 + kk) * ny1 + jj; };
  for (int j = 0; j < k; j++) {
    traction_face(P, seed, case0 + j, t.data());
    double* fj = f + (size_t)j * np;
    int64_t off = 0;
    for (const Subdomain& s : P.subs) {
      std::fill(fj + off, fj + off + s.n, 0.0);
      if (s.e0[0] + s.ne[0] == NX) {
        const int ii = s.nn[0] - 1;
        for (int b = 0; b < s.nn[2]; b++)
          for (int a = 0; a < s.nn[1]; a++) {
            const int c1 = (a > 0) + (a < s.ne[1]), c2 = (b > 0) + (b < s.ne[2]);
            const double w = (double)(c1 * c2) * hy * hz * 0.25;
            const int ln = ii + a * s.nn[0] + b * s.nn[0] * s.nn[1];
            for (int c = 0; c < 3; c++) fj[off + 3 * ln + c] = t[at(c, s.e0[1] + a, s.e0[2] + b)] * w;
          }
      }
      off += s.n;
    }
  }
}

}  // namespace hfls
