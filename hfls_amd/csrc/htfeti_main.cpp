// Drop-in check of the C++ facade: the reference's main (htfeti.cpp:8-110) with
// the same options map and positional Nx Ny Nz nx ny nz arguments, running the
// device path through hfls::Driver.
#include <cstdio>
#include <map>
#include <string>

#include "hfls/htfeti.hpp"

int main(int argc, char* argv[]) {
  std::map<std::string, std::string> options2;
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
  try {
    hfls::Driver driver(options2);
    driver.initialization();
    const hfls::Cluster& c = driver.cluster;
    double umax = 0;
    for (double v : c.solution) umax = v < -umax ? -v : (v > umax ? v : umax);
    std::printf("iterations %d converged %d residual %.3e |Pg|/|Pg0| %.3e max|u| %.6e\n", c.stats.iterations,
                c.stats.converged, c.stats.residual, c.stats.pg_rel, umax);
    // operator round trip through the facade: F is symmetric
    hfls::Cluster& cc = driver.cluster;
    std::vector<double> a(cc.get_nLam_f()), b(cc.get_nLam_f()), Fa, Fb;
    for (int i = 0; i < cc.get_nLam_f(); i++) { a[i] = (i % 7) - 3.0; b[i] = (i % 5) - 2.0; }
    cc.mult_Ff(a, Fa);
    cc.mult_Ff(b, Fb);
    double ab = 0, ba = 0;
    for (int i = 0; i < cc.get_nLam_f(); i++) { ab += a[i] * Fb[i]; ba += b[i] * Fa[i]; }
    std::printf("symmetry |a'Fb - b'Fa|/|a'Fb| = %.3e\n", (ab - ba) / (ab != 0 ? ab : 1));
    return c.stats.converged ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
}
