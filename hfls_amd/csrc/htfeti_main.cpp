// Drop-in check of the C++ facade: the reference's main (htfeti.cpp:8-110) running
// the device path through hfls::Driver.
//
//   htfeti [Nx Ny Nz nx ny nz]           x = 0 clamped cube, body force, 1e-8
//   htfeti --reference-map <dir> [...]   the reference's own options map
//                                        (htfeti.cpp:12-57) unchanged except
//                                        path2data = <dir>; the positional
//                                        Nx Ny Nz nx ny nz nparts follow
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "hfls/htfeti.hpp"

int main(int argc, char* argv[]) {
  std::map<std::string, std::string> options2;
  int first = 1;
  if (argc > 2 && std::strcmp(argv[1], "--reference-map") == 0) {
    // htfeti.cpp:12-57, value for value
    options2["eps_iter"] = "1e-6";
    options2["max_iter"] = "200";
    options2["vtkWithinIter"] = "true";
    options2["metis"] = "true";
    options2["preconditioner"] = "Dirichlet_implicit";
    options2["Ac_clust_explicit"] = "true";
    options2["path2data"] = argv[2];
    options2["young_modulus"] = "1000";
    options2["ratio_mat"] = "1e4";
    options2["poissons_ratio"] = "0.3";
    options2["linear_solver"] = "pardiso";
    options2["print_matrices"] = "1";
    options2["typeBc"] = "ker";
    options2["Ac_extended_by_kerGc"] = "false";
    options2["GcTGc_assembl_block_by_block"] = "true";
    options2["Bc_fullRank"] = "true";
    options2["Nx"] = "2";
    options2["Ny"] = "1";
    options2["Nz"] = "1";
    options2["nx"] = "5";
    options2["ny"] = "5";
    options2["nz"] = "5";
    options2["nparts"] = "1";
    if (options2["linear_solver"] == "pardiso" && options2["create_analytic_ker_K"] == "false")
      options2["create_analytic_ker_K"] = "true";
    first = 3;
    const char* keys[] = {"Nx", "Ny", "Nz", "nx", "ny", "nz", "nparts"};
    for (int i = 0; i < 7 && first + i < argc; i++) options2[keys[i]] = argv[first + i];
  } else {
    options2["eps_iter"] = "1e-8";
    options2["max_iter"] = "200";
    options2["preconditioner"] = "Dirichlet_explicit";
    options2["typeBc"] = "ker";
    options2["bc"] = "clamp_x0";
    options2["material"] = "uniform";
    options2["young_modulus"] = "2.1e5";
    options2["load"] = "body";
    const char* keys[] = {"Nx", "Ny", "Nz", "nx", "ny", "nz"};
    const char* defs[] = {"2", "2", "2", "3", "3", "3"};
    for (int i = 0; i < 6; i++) options2[keys[i]] = argc > i + 1 ? argv[i + 1] : defs[i];
  }
  try {
    hfls::Driver driver(options2);
    driver.initialization();
    hfls::Cluster& c = driver.cluster;
    double umax = 0;
    for (double v : c.solution) umax = v < -umax ? -v : (v > umax ? v : umax);
    std::printf("subdomains %d nLam_f %d nLam_c %d nRBM_f %d nRBM_c %d neqClst %d\n", c.get_nSubClst(),
                c.get_nLam_f(), c.get_nLam_c(), c.get_nRBM_f(), c.get_nRBM_c(), c.get_neqClst());
    std::printf("iterations %d converged %d residual %.3e |Pg|/|Pg0| %.3e max|u| %.9e\n", c.stats.iterations,
                c.stats.converged, c.stats.residual, c.stats.pg_rel, umax);
    // operator round trips through the facade: F is symmetric, P = I - Gf (Gf^T Gf)^-1 Gf^T
    // annihilates Gf, and the recovery of printVTK from (lambda, alpha of P g) is the solution
    std::vector<double> a(c.get_nLam_f()), b(c.get_nLam_f()), Fa, Fb;
    for (int i = 0; i < c.get_nLam_f(); i++) { a[i] = (i % 7) - 3.0; b[i] = (i % 5) - 2.0; }
    c.mult_Ff(a, Fa);
    c.mult_Ff(b, Fb);
    double ab = 0, ba = 0;
    for (int i = 0; i < c.get_nLam_f(); i++) { ab += a[i] * Fb[i]; ba += b[i] * Fa[i]; }
    std::printf("symmetry |a'Fb - b'Fa|/|a'Fb| = %.3e\n", (ab - ba) / (ab != 0 ? ab : 1));
    std::vector<double> al(c.get_nRBM_f()), Ga, PGa, beta;
    for (int i = 0; i < c.get_nRBM_f(); i++) al[i] = 1.0 + i;
    c.mult_Gf(al, Ga);
    c.Projection(Ga, PGa, beta);
    double ng = 0, npg = 0;
    for (int i = 0; i < c.get_nLam_f(); i++) { ng += Ga[i] * Ga[i]; npg += PGa[i] * PGa[i]; }
    std::printf("projector |P Gf a|/|Gf a| = %.3e\n", ng > 0 ? std::sqrt(npg / ng) : 0.0);
    return c.stats.converged ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
}
