// Drop-in C++ facade of the reference's Driver / Mesh / Cluster / Solver entry
// points (Driver.hpp:11-24, Mesh.hpp:60-87, Cluster.hpp:15-138, Solver.hpp:12-20)
// over the C ABI of libhfls.so.  Same class and method names, same
// map<string,string> options (htfeti.cpp:12-57), same operator semantics; vectors
// are std::vector<double> (the reference's Vector is a column-major dense Matrix).
//
// A Cluster keeps every factorisation and explicit operator on the GPU; the
// methods move dual vectors (nLam_f doubles, reference Bf row order) and
// per-subdomain vectors (l2g order, Data.cpp:111-144) between host and device.
#pragma once
#include <cmath>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hfls_gpu.h"

namespace hfls {

typedef std::vector<double> Vector;
typedef std::map<std::string, std::string> OptionMap;

inline void check(int rc) {
  if (rc != 0) throw std::runtime_error(hfls_last_error());
}

struct KV {
  std::vector<const char*> k, v;
  explicit KV(const OptionMap& o) {
    for (auto& e : o) {
      k.push_back(e.first.c_str());
      v.push_back(e.second.c_str());
    }
  }
};

// Mesh::createMesh (Mesh.cpp:69-289).  Mesh::ddm_metis (Mesh.cpp:352-452) is
// folded into createMesh: metis=false keeps the geometric Nx x Ny x Nz partition,
// metis=true with nparts <= 1 makes the whole cube one subdomain (what ddm_metis
// does without calling METIS, Mesh.cpp:413-421), metis=true with nparts > 1 is an
// error (METIS is not available).
class Mesh {
 public:
  Mesh() = default;
  ~Mesh() { if (p) hfls_problem_destroy(p); }
  Mesh(const Mesh&) = delete;
  Mesh& operator=(const Mesh&) = delete;
  void createMesh(OptionMap& o) {
    KV kv(o);
    if (p) hfls_problem_destroy(p);
    p = nullptr;
    check(hfls_problem_create((int)kv.k.size(), kv.k.data(), kv.v.data(), &p));
    int64_t s[8];
    check(hfls_problem_sizes(p, s));
    nSubClst = (int)s[0];
    nPoints = (int)s[3];
  }
  void ddm_metis(OptionMap&) {}
  int nSubClst = 0, nPoints = 0;
  hfls_problem handle() const { return p; }

 private:
  hfls_problem p = nullptr;
};

class Cluster {
 public:
  Cluster() = default;
  ~Cluster() { if (m) hfls_mesh_destroy(m); }
  Cluster(const Cluster&) = delete;
  Cluster& operator=(const Cluster&) = delete;

  // Cluster::initialization (Cluster.cpp:21-385): setup on the GPU, then pcpg()
  void initialization(OptionMap o, Mesh& mesh, bool run_pcpg = true) {
    options2 = o;
    mesh_ = &mesh;
    KV kv(o);
    int dev = o.count("device") ? std::atoi(o["device"].c_str()) : 0;
    check(hfls_mesh_create(dev, (int)kv.k.size(), kv.k.data(), kv.v.data(), &m));
    int64_t s[6];
    check(hfls_mesh_sizes(m, s));
    nSubClst = (int)s[0];
    nLam_f = (int)s[1];
    nRBM_f = (int)s[2];
    nPacked = s[3];
    n_sub.resize(nSubClst);
    for (int d = 0; d < nSubClst; d++) {
      int info[8];
      check(hfls_problem_sub_info(mesh.handle(), d, info));
      n_sub[d] = info[0];
    }
    std::vector<double> w(nLam_f);
    std::vector<int> a(nLam_f), b(nLam_f);
    check(hfls_problem_rows(mesh.handle(), a.data(), b.data(), nullptr, w.data()));
    weigth = w;
    if (run_pcpg) pcpg();
  }

  // Cluster.hpp:26-31
  int get_nSubClst() { return nSubClst; }
  int get_neqClst() { return (int)counts(1); }
  int get_nLam_c() { return (int)counts(2); }
  int get_nLam_f() { return nLam_f; }
  int get_nRBM_f() { return nRBM_f; }
  int get_nRBM_c() { return (int)counts(5); }

  // lambda = Gf alpha (Cluster.cpp:1275-1283)
  void mult_Gf(Vector const& alpha, Vector& lambda) {
    if ((int)alpha.size() != nRBM_f) throw std::runtime_error("alpha has the wrong size");
    lambda.resize(nLam_f);
    check(hfls_mult_Gf(m, alpha.data(), lambda.data()));
  }
  // alpha = Gf^T lambda (Cluster.cpp:1285-1293)
  void mult_GfT(Vector const& lambda, Vector& alpha) {
    if ((int)lambda.size() != nLam_f) throw std::runtime_error("lambda has the wrong size");
    alpha.resize(nRBM_f);
    check(hfls_mult_GfT(m, lambda.data(), alpha.data()));
  }
  // alpha = sum_d Rf_d^T x_d (Cluster.cpp:1295-1312)
  void mult_RfT(std::vector<Vector> const& x_in, Vector& alpha) {
    Vector f = pack(x_in);
    alpha.resize(nRBM_f);
    check(hfls_mult_RfT(m, f.data(), alpha.data()));
  }

  // x = K_c^+ f (Cluster.cpp:967-1040)
  void mult_Kplus_f(std::vector<Vector>& rhs_in, std::vector<Vector>& x_out) {
    Vector f = pack(rhs_in), x(nPacked);
    check(hfls_kplus(m, f.data(), x.data()));
    x_out = unpack(x);
  }
  // F w = Bf K_c^+ Bf^T w (Cluster.cpp:1114-1188)
  void mult_Ff(Vector const& w, Vector& Fw) {
    Fw.resize(nLam_f);
    check(hfls_apply_F(m, w.data(), Fw.data()));
  }
  // M x = W Bf S Bf^T W x (Cluster.cpp:1043-1112)
  void Preconditioning(Vector& w_in, Vector& w_out) {
    w_out.resize(nLam_f);
    check(hfls_precondition(m, w_in.data(), w_out.data()));
  }
  // Px = x - Gf (GfTGf)^-1 Gf^T x; alpha = -(GfTGf)^-1 Gf^T x (Cluster.cpp:1314-1329)
  void Projection(Vector& x, Vector& Px, Vector& alpha) {
    Px.resize(nLam_f);
    alpha.resize(nRBM_f);
    check(hfls_project(m, x.data(), Px.data(), alpha.data()));
  }
  // lambda = sum_d Bf_d x_d (Cluster.cpp:1193-1208)
  void mult_Bf(std::vector<Vector> const& x_in, Vector& lambda) {
    lambda.assign(nLam_f, 0.0);
    for (int d = 0; d < nSubClst; d++) {
      std::vector<int> r, c, s;
      bf(d, r, c, s);
      for (size_t e = 0; e < r.size(); e++) lambda[r[e]] += s[e] * x_in[d][c[e]];
    }
  }
  // x_d = Bf_d^T lambda (Cluster.cpp:1210-1224)
  void mult_BfT(Vector const& lambda, std::vector<Vector>& x_out) {
    x_out.resize(nSubClst);
    for (int d = 0; d < nSubClst; d++) {
      x_out[d].assign(n_sub[d], 0.0);
      std::vector<int> r, c, s;
      bf(d, r, c, s);
      for (size_t e = 0; e < r.size(); e++) x_out[d][c[e]] += s[e] * lambda[r[e]];
    }
  }
  // multiplicity scaling (Cluster.cpp:1356-1360)
  void scale(Vector& x) {
    for (int i = 0; i < nLam_f; i++) x[i] *= weigth[i];
  }
  // Cluster::pcpg (Cluster.cpp:1363-1458) and the final recovery of printVTK; with
  // path2data set the library writes <path2data>/box.vtk, and with
  // vtkWithinIter=true also box<it>.vtk after every iteration (Cluster.cpp:1448-1456)
  void pcpg() {
    double eps = options2.count("eps_iter") ? std::atof(options2["eps_iter"].c_str()) : 1e-6;
    int max_iter = options2.count("max_iter") ? std::atoi(options2["max_iter"].c_str()) : 200;
    int64_t s[6];
    check(hfls_mesh_sizes(m, s));
    solution.assign(3 * s[4], 0.0);
    lambda.assign(nLam_f, 0.0);
    check(hfls_solve(m, eps, max_iter, nullptr, solution.data(), lambda.data(), &stats));
  }
  // Cluster::printVTK (Cluster.cpp:1648-1678): yy = rhs - Bf^T lambda, xx = K_c^+ yy -
  // Rf alpha per subdomain, the scatter of xx written as <path2data>/box.vtk (it < 0)
  // or box<it>.vtk; rhs is the options' body force
  void printVTK(std::vector<Vector>& yy, std::vector<Vector>& xx, Vector& lambda_in, Vector& alpha, int it) {
    if ((int)lambda_in.size() != nLam_f || (int)alpha.size() != nRBM_f)
      throw std::runtime_error("lambda or alpha has the wrong size");
    Vector f(nPacked), x(nPacked);
    check(hfls_mesh_load(m, f.data()));
    std::vector<Vector> bt;
    mult_BfT(lambda_in, bt);
    yy = unpack(f);
    for (int d = 0; d < nSubClst; d++)
      for (size_t i = 0; i < yy[d].size(); i++) yy[d][i] -= bt[d][i];
    int64_t s[6];
    check(hfls_mesh_sizes(m, s));
    Vector u(3 * s[4]);
    check(hfls_recover(m, f.data(), lambda_in.data(), alpha.data(), x.data(), u.data()));
    xx = unpack(x);
    check(hfls_mesh_write_vtk(m, u.data(), options2["path2data"].c_str(), it));
  }
  // the displacements of the last pcpg() as VTK (<path2data>/box.vtk for it < 0)
  void printVTK(int it) {
    check(hfls_mesh_write_vtk(m, solution.data(), options2["path2data"].c_str(), it));
  }

  OptionMap options2;
  std::vector<double> weigth;  // reference spelling (Cluster.hpp:95)
  Vector solution, lambda;
  hfls_stats stats{};
  hfls_mesh handle() const { return m; }

 private:
  hfls_mesh m = nullptr;
  Mesh* mesh_ = nullptr;
  int nSubClst = 0, nLam_f = 0, nRBM_f = 0;
  int64_t nPacked = 0;
  std::vector<int> n_sub;
  int64_t counts(int i) {
    int64_t c[6];
    check(hfls_mesh_counts(m, c));
    return c[i];
  }
  Vector pack(const std::vector<Vector>& v) {
    Vector out;
    out.reserve(nPacked);
    for (auto& x : v) out.insert(out.end(), x.begin(), x.end());
    if ((int64_t)out.size() != nPacked) throw std::runtime_error("per-subdomain vector sizes do not match");
    return out;
  }
  std::vector<Vector> unpack(const Vector& x) {
    std::vector<Vector> out(nSubClst);
    size_t o = 0;
    for (int d = 0; d < nSubClst; d++) {
      out[d].assign(x.begin() + o, x.begin() + o + n_sub[d]);
      o += n_sub[d];
    }
    return out;
  }
  void bf(int d, std::vector<int>& r, std::vector<int>& c, std::vector<int>& s) {
    int info[8];
    check(hfls_problem_sub_info(mesh_->handle(), d, info));
    r.resize(info[6]);
    c.resize(info[6]);
    s.resize(info[6]);
    check(hfls_problem_sub_bf(mesh_->handle(), d, r.data(), c.data(), s.data()));
  }
};

// Solver::pcpg(map, Cluster&) (Solver.cpp:11-105) runs the same PCPG as
// Cluster::pcpg on the cluster's device state.
class Solver {
 public:
  void pcpg(OptionMap o, Cluster& c) {
    for (auto& e : o) c.options2[e.first] = e.second;
    c.pcpg();
  }
};

// Driver (Driver.cpp:14-38): createMesh -> ddm_metis -> cluster.initialization
class Driver {
 public:
  Driver() = default;
  explicit Driver(OptionMap o) : options2(o) {}
  void initialization() {
    mesh.createMesh(options2);
    mesh.ddm_metis(options2);
    cluster.initialization(options2, mesh);
  }
  Mesh mesh;
  Cluster cluster;
  Solver solver;

 private:
  OptionMap options2;
};

}  // namespace hfls
