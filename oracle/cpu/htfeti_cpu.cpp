// HTFETI on the CPU in FP64 with OpenMP over subdomains.  TEST INFRASTRUCTURE
// ONLY: the CPU baseline that bench.py times and tests/ check against the Python
// oracle; nothing in hfls_amd/ loads it.
//
// It restates the reference's PARDISO path (the one its main() selects,
// htfeti.cpp:27-62: linear_solver=pardiso forces create_analytic_ker_K=true,
// preconditioner=Dirichlet_implicit):
//   * hex8 element stiffness and body load, 2x2x2 Gauss points (Data.cpp:297-490),
//     assembled per subdomain (Data.cpp:25-54, 231-273);
//   * analytic rigid-body kernel R_d from the global node coordinates, columns
//     normalised (Data.cpp:61-109);
//   * fixing DOFs by full-pivot elimination on R (Matrix::getNullPivots,
//     Matrix.cpp:835-884), K_reg = K with those diagonals doubled
//     (Matrix::num_factor, Matrix.cpp:1029-1060), factorised once;
//   * Dirichlet_implicit: K_rr factorised, S x = K_bb x - K_br K_rr^-1 K_rb x
//     applied per iteration (Matrix.cpp:1853-1870, Cluster.cpp:1090-1103);
//   * HTFETI cluster objects KplusBcT, Fc, Gc (Cluster.cpp:151-206), the kernel H
//     of Gc^T Gc and the extended cluster matrix [Fc Gc 0; Gc^T 0 H; 0 H^T 0]
//     (Ac_extended_by_kerGc=true, Cluster.cpp:592-660), LU-factorised;
//   * Rf = R H, Gf = Bf Rf, (Gf^T Gf)^-1 (Cluster.cpp:1227-1273, 1332-1352);
//   * PCPG with an implicit K_c^+ per iteration (Cluster.cpp:1363-1458, mult_Ff
//     1114-1188, mult_Kplus_f 967-1040, Projection 1314-1329, Preconditioning
//     1043-1112) and the recovery of printVTK (Cluster.cpp:1648-1678).
// PARDISO's sparse LDL^T is replaced by a band Cholesky in the l2g order (on the
// structured cube the half bandwidth is 3((nx+1)(ny+1) + nx + 2) - 1); H is the
// global rigid-body motion written in each subdomain's R basis (the kernel of Gc
// for any gluing that the global rigid motions satisfy), orthonormalised, instead
// of the reference's numerical kernel detection on Gc^T Gc; P F P and u do not
// depend on the basis of H.
//
// Reductions over dual rows use a fixed number of chunks added in order, so the
// results do not depend on the thread count.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); }

// ---------------------------------------------------------------- band Cholesky
// Lower band storage: A(i, j), j <= i <= j + kd, at ab[(i - j) + j (kd + 1)].
struct Band {
  int n = 0, kd = 0;
  std::vector<double> ab;
  double& at(int i, int j) { return ab[(size_t)(i - j) + (size_t)j * (kd + 1)]; }
  // in place A = L L^T, right-looking; the update of column j + k is contiguous
  void factor() {
    const int ld = kd + 1;
    for (int j = 0; j < n; j++) {
      double* cj = ab.data() + (size_t)j * ld;
      const int m = std::min(kd, n - 1 - j);
      if (!(cj[0] > 0.0)) throw std::runtime_error("band Cholesky: non-positive pivot");
      const double d = std::sqrt(cj[0]), inv = 1.0 / d;
      cj[0] = d;
      for (int i = 1; i <= m; i++) cj[i] *= inv;
      for (int k = 1; k <= m; k++) {
        double* ck = ab.data() + (size_t)(j + k) * ld;
        const double l = cj[k];
        const double* src = cj + k;
        const int len = m - k + 1;
        for (int i = 0; i < len; i++) ck[i] -= l * src[i];
      }
    }
  }
  // x <- (L L^T)^-1 x
  void solve(double* x) const {
    const int ld = kd + 1;
    for (int j = 0; j < n; j++) {
      const double* cj = ab.data() + (size_t)j * ld;
      const int m = std::min(kd, n - 1 - j);
      const double v = x[j] / cj[0];
      x[j] = v;
      double* xj = x + j;
      for (int i = 1; i <= m; i++) xj[i] -= cj[i] * v;
    }
    for (int j = n - 1; j >= 0; j--) {
      const double* cj = ab.data() + (size_t)j * ld;
      const int m = std::min(kd, n - 1 - j);
      const double* xj = x + j;
      double s = 0.0;
      for (int i = 1; i <= m; i++) s += cj[i] * xj[i];
      x[j] = (x[j] - s) / cj[0];
    }
  }
  // X (n x r, row major) <- (L L^T)^-1 X
  void solve_rows(double* X, int r) const {
    const int ld = kd + 1;
    for (int j = 0; j < n; j++) {
      const double* cj = ab.data() + (size_t)j * ld;
      const int m = std::min(kd, n - 1 - j);
      double* xj = X + (size_t)j * r;
      const double inv = 1.0 / cj[0];
      for (int c = 0; c < r; c++) xj[c] *= inv;
      for (int i = 1; i <= m; i++) {
        double* xi = X + (size_t)(j + i) * r;
        const double l = cj[i];
        for (int c = 0; c < r; c++) xi[c] -= l * xj[c];
      }
    }
    for (int j = n - 1; j >= 0; j--) {
      const double* cj = ab.data() + (size_t)j * ld;
      const int m = std::min(kd, n - 1 - j);
      double* xj = X + (size_t)j * r;
      for (int i = 1; i <= m; i++) {
        const double* xi = X + (size_t)(j + i) * r;
        const double l = cj[i];
        for (int c = 0; c < r; c++) xj[c] -= l * xi[c];
      }
      const double inv = 1.0 / cj[0];
      for (int c = 0; c < r; c++) xj[c] *= inv;
    }
  }
};

// ---------------------------------------------------------------- CSR (general)
struct Csr {
  int rows = 0;
  std::vector<int> ptr, col;
  std::vector<double> val;
  // y (+)= A x
  void apply(const double* x, double* y, bool add) const {
    for (int i = 0; i < rows; i++) {
      double s = 0.0;
      for (int p = ptr[i]; p < ptr[i + 1]; p++) s += val[p] * x[col[p]];
      y[i] = add ? y[i] + s : s;
    }
  }
};

// ---------------------------------------------------------------- dense LU
// column-major n x n, partial pivoting (the cluster saddle-point matrix)
struct Lu {
  int n = 0;
  std::vector<double> a;
  std::vector<int> piv;
  void factor() {
    piv.resize(n);
    for (int j = 0; j < n; j++) {
      double* cj = a.data() + (size_t)j * n;
      int p = j;
      double best = std::fabs(cj[j]);
      for (int i = j + 1; i < n; i++)
        if (std::fabs(cj[i]) > best) best = std::fabs(cj[i]), p = i;
      if (!(best > 0.0)) throw std::runtime_error("cluster matrix singular");
      piv[j] = p;
      if (p != j)
        for (int k = 0; k < n; k++) std::swap(a[(size_t)k * n + j], a[(size_t)k * n + p]);
      const double inv = 1.0 / cj[j];
      for (int i = j + 1; i < n; i++) cj[i] *= inv;
      for (int k = j + 1; k < n; k++) {
        double* ck = a.data() + (size_t)k * n;
        const double u = ck[j];
        if (u == 0.0) continue;
        for (int i = j + 1; i < n; i++) ck[i] -= cj[i] * u;
      }
    }
  }
  void solve(double* x) const {
    for (int j = 0; j < n; j++)
      if (piv[j] != j) std::swap(x[j], x[piv[j]]);
    for (int j = 0; j < n; j++) {
      const double* cj = a.data() + (size_t)j * n;
      const double v = x[j];
      for (int i = j + 1; i < n; i++) x[i] -= cj[i] * v;
    }
    for (int j = n - 1; j >= 0; j--) {
      const double* cj = a.data() + (size_t)j * n;
      x[j] /= cj[j];
      const double v = x[j];
      for (int i = 0; i < j; i++) x[i] -= cj[i] * v;
    }
  }
};

// ---------------------------------------------------------------- hex8
// natural coordinates of the 8 nodes in the reference's element node order
// (Mesh.cpp:247-263): (i,j),(i+1,j),(i+1,j+1),(i,j+1) at k, then at k+1
const double XI[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
const double ETA[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
const double ZETA[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// Ke (24 x 24, DOF 3a + c, row major) and fe (24) of one element: the isoparametric
// trilinear brick, isotropic C in Voigt order xx, yy, zz, xy, yz, xz, 2x2x2 Gauss
// points (weights 1), fe = acc N detJ
void hex8(const double* X /*8 x 3*/, double E, double nu, const double* acc, double* Ke, double* fe) {
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
  const double g = 1.0 / std::sqrt(3.0);
  std::memset(Ke, 0, sizeof(double) * 576);
  std::memset(fe, 0, sizeof(double) * 24);
  for (int gp = 0; gp < 8; gp++) {
    const double r = (gp & 1) ? g : -g, s = (gp & 2) ? g : -g, t = (gp & 4) ? g : -g;
    double dN[3][8], N[8];
    for (int a = 0; a < 8; a++) {
      const double pr = 1 + r * XI[a], ps = 1 + s * ETA[a], pt = 1 + t * ZETA[a];
      N[a] = pr * ps * pt / 8;
      dN[0][a] = XI[a] * ps * pt / 8;
      dN[1][a] = ETA[a] * pr * pt / 8;
      dN[2][a] = ZETA[a] * pr * ps / 8;
    }
    double J[3][3] = {};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        for (int a = 0; a < 8; a++) J[i][j] += dN[i][a] * X[3 * a + j];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double Ji[3][3];
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    double dx[8][3];  // dN_a / dx_j = sum_i Ji[j][i] dN[i][a]
    for (int a = 0; a < 8; a++)
      for (int j = 0; j < 3; j++) dx[a][j] = Ji[j][0] * dN[0][a] + Ji[j][1] * dN[1][a] + Ji[j][2] * dN[2][a];
    double B[6][24] = {};
    for (int a = 0; a < 8; a++) {
      B[0][3 * a] = dx[a][0];
      B[1][3 * a + 1] = dx[a][1];
      B[2][3 * a + 2] = dx[a][2];
      B[3][3 * a] = dx[a][1], B[3][3 * a + 1] = dx[a][0];
      B[4][3 * a + 1] = dx[a][2], B[4][3 * a + 2] = dx[a][1];
      B[5][3 * a] = dx[a][2], B[5][3 * a + 2] = dx[a][0];
    }
    double D[6][24];  // C B
    for (int q = 0; q < 24; q++) {
      const double e0 = B[0][q], e1 = B[1][q], e2 = B[2][q];
      D[0][q] = (lam + 2 * mu) * e0 + lam * (e1 + e2);
      D[1][q] = (lam + 2 * mu) * e1 + lam * (e0 + e2);
      D[2][q] = (lam + 2 * mu) * e2 + lam * (e0 + e1);
      D[3][q] = mu * B[3][q];
      D[4][q] = mu * B[4][q];
      D[5][q] = mu * B[5][q];
    }
    for (int p = 0; p < 24; p++)
      for (int v = 0; v < 6; v++) {
        const double b = B[v][p] * det;
        if (b == 0.0) continue;
        for (int q = 0; q < 24; q++) Ke[p * 24 + q] += b * D[v][q];
      }
    for (int a = 0; a < 8; a++)
      for (int c = 0; c < 3; c++) fe[3 * a + c] += N[a] * acc[c] * det;
  }
}

// Matrix::getNullPivots (Matrix.cpp:835-884) on R (n x 6, column major): full-pivot
// elimination, the pivot found as the last maximum of |N| in column-major storage
// over the leading columns, swapped to the bottom rows; returns the sorted pivots
std::vector<int> null_pivots(const std::vector<double>& R, int n, int nc) {
  std::vector<double> N(R);
  std::vector<int> piv(n);
  for (int i = 0; i < n; i++) piv[i] = i;
  auto ij = [n](int i, int j) { return (size_t)i + (size_t)j * n; };
  std::vector<double> tmp(n);
  for (int j = 0; j < nc; j++) {
    const size_t len = (size_t)n * (nc - j);
    size_t I = 0;
    double best = -1.0;
    for (size_t q = 0; q < len; q++)
      if (std::fabs(N[q]) >= best) best = std::fabs(N[q]), I = q;
    const int colInd = (int)(I / n), rowInd = (int)(I % n);
    const int rlast = n - 1 - j, clast = nc - 1 - j;
    for (int k = 0; k < nc - j; k++) std::swap(N[ij(rlast, k)], N[ij(rowInd, k)]);
    std::swap(piv[rowInd], piv[rlast]);
    if (clast != colInd)
      for (int i = 0; i < n; i++) std::swap(N[ij(i, clast)], N[ij(i, colInd)]);
    const double pivot = N[ij(rlast, clast)];
    for (int J = 0; J < clast; J++) {  // (a b) / pivot in the reference's order: ties depend on it
      const double b = N[ij(rlast, J)];
      for (int i = 0; i < n - j; i++) N[ij(i, J)] -= N[ij(i, clast)] * b / pivot;
    }
  }
  std::vector<int> out(nc);
  for (int i = 0; i < nc; i++) out[i] = piv[n - 1 - i];
  std::sort(out.begin(), out.end());
  return out;
}

constexpr int NCHUNK = 256;  // fixed reduction chunks over dual rows

// body(i) for i < n on the OpenMP threads (one index per dynamic chunk); an
// exception of a body is rethrown after the loop, since none may leave an OpenMP
// region
template <class F>
void par_for(int n, F body) {
  bool bad = false;
  std::string err;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < n; i++) {
    try {
      body(i);
    } catch (const std::exception& e) {
#pragma omp critical(htc_err)
      if (!bad) bad = true, err = e.what();
    }
  }
  if (bad) throw std::runtime_error(err);
}

}  // namespace

extern "C" {

// Problem description (packed per-subdomain arrays; local DOF = rank in l2g)
struct htc_desc {
  int nsub, ncl;
  int64_t nlam, ndof;      // dual rows; global DOFs (3 nPoints)
  double nu, acc[3];       // Poisson ratio; body load per unit volume
  const int64_t* n_off;    // [nsub + 1] DOF offsets
  const int64_t* l2g;      // [n_off[nsub]] global DOF of each local DOF
  const double* coords;    // [n_off[nsub]] node coordinates, (x, y, z) at local DOFs 3a .. 3a + 2
  const int64_t* e_off;    // [nsub + 1] element offsets
  const int32_t* conn;     // [e_off[nsub] * 8] local node ids
  const double* Eel;       // [e_off[nsub]] Young's modulus per element
  const int64_t* bf_off;   // [nsub + 1] Bf entry offsets
  const int64_t* bf_row;   // dual row of each entry
  const int32_t* bf_col;   // local DOF
  const double* bf_val;    // +-1
  const double* weight;    // [nlam] preconditioner scaling
  const int64_t* bc_off;   // [nsub + 1] Bc entry offsets
  const int32_t* bc_row;   // cluster-local Bc row
  const int32_t* bc_col;   // local DOF
  const double* bc_val;
  const int32_t* cl_d0;    // [ncl] first subdomain of each cluster
  const int32_t* cl_d1;    // [ncl] one past the last
  const int32_t* cl_nlc;   // [ncl] Bc rows of each cluster
  int lumped;              // 0: Dirichlet_implicit, 1: lumped (K)
};

struct htc_times {
  double assemble, kernel_factor, dirichlet, cluster, coarse, total;
};
struct htc_solve_times {
  double init, pcpg, recover, total;
};

}  // extern "C"

namespace {

struct Sub {
  int n = 0, cl = 0;
  Band K;                      // K_reg, then its factor
  Csr Kfull;                   // K (both triangles), for the lumped preconditioner
  std::vector<double> R;       // n x 6, column major (analytic, normalised)
  std::vector<double> f;       // body load
  std::vector<int> fix;        // fixing DOFs
  std::vector<int> bdofs;      // j_col_cmpr: sorted B-touched local DOFs
  std::vector<int> rdofs;      // the others
  Band Krr;                    // interior factor (Dirichlet_implicit)
  Csr Krb, Kbr, Kbb;           // blocks of K in (r, b) numbering
  std::vector<int> bc_rows;    // sorted unique cluster-local Bc rows
  std::vector<double> KBcT;    // n x nbc (row major): K_reg^-1 Bc^T
  std::vector<double> BcT;     // n x nbc (row major)
  std::vector<double> Gc;      // nbc x 6 (row major): -Bc R
  int roff = 0;                // offset of alpha_d in the cluster's alpha block
  std::vector<double> RH;      // n x 6 (column major): R H_d, this cluster's 6 Gf columns
};

struct Cl {
  int d0, d1, nlc, nr, na;
  Lu Ac;
  std::vector<double> H;  // nr x 6, row major
};

struct Ent {
  int d, col;
  double val;
};

struct Problem {
  htc_desc D;
  std::vector<Sub> S;
  std::vector<Cl> C;
  int64_t nlam = 0;
  int ncol = 0;
  std::vector<int64_t> rptr;       // dual row -> entries
  std::vector<Ent> rent;
  std::vector<double> Grow;        // per entry 6 values of (Bf R H) for the entry's cluster
  std::vector<double> invGtG;      // ncol x ncol
  std::vector<double> weight;
};

// ------------------------------------------------------------ dual helpers
double dot(const Problem& P, const double* a, const double* b) {
  double part[NCHUNK];
  const int64_t n = P.nlam;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < NCHUNK; c++) {
    const int64_t r0 = n * c / NCHUNK, r1 = n * (c + 1) / NCHUNK;
    double s = 0.0;
    for (int64_t r = r0; r < r1; r++) s += a[r] * b[r];
    part[c] = s;
  }
  double s = 0.0;
  for (int c = 0; c < NCHUNK; c++) s += part[c];
  return s;
}

// x_d = Bf_d^T lam (Cluster.cpp:1210-1224)
void mult_BfT(const Problem& P, const double* lam, std::vector<std::vector<double>>& xs, const double* scale) {
  const int N = P.D.nsub;
  par_for(N, [&](int d) {
    std::vector<double>& x = xs[d];
    std::fill(x.begin(), x.end(), 0.0);
    for (int64_t p = P.D.bf_off[d]; p < P.D.bf_off[d + 1]; p++) {
      const int64_t r = P.D.bf_row[p];
      x[P.D.bf_col[p]] += P.D.bf_val[p] * lam[r] * (scale ? scale[r] : 1.0);
    }
  });
}

// lam = sum_d Bf_d x_d (Cluster.cpp:1193-1208), per dual row in subdomain order
void mult_Bf(const Problem& P, const std::vector<std::vector<double>>& xs, double* lam, const double* scale) {
  const int64_t n = P.nlam;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    double s = 0.0;
    for (int64_t e = P.rptr[r]; e < P.rptr[r + 1]; e++) s += P.rent[e].val * xs[P.rent[e].d][P.rent[e].col];
    lam[r] = scale ? s * scale[r] : s;
  }
}

// K_c^+ f (Cluster.cpp:967-1040): x_d = K_reg^-1 f_d + R_d alpha_d - KBcT_d lam_c with
// [lam_c; alpha; mu] = Ac^-1 [KBcT^T f; -R^T f; 0]
void mult_Kplus_f(const Problem& P, const std::vector<std::vector<double>>& fs, std::vector<std::vector<double>>& xs) {
  const int N = P.D.nsub, NC = P.D.ncl;
  std::vector<std::vector<double>> la(NC);
  par_for(NC, [&](int c) {
    const Cl& cl = P.C[c];
    std::vector<double> gc(cl.na, 0.0);
    for (int d = cl.d0; d < cl.d1; d++) {
      const Sub& s = P.S[d];
      const int nb = (int)s.bc_rows.size();
      const double* f = fs[d].data();
      for (int q = 0; q < nb; q++) {
        double t = 0.0;
        for (int i = 0; i < s.n; i++) t += s.KBcT[(size_t)i * nb + q] * f[i];
        gc[s.bc_rows[q]] += t;
      }
      for (int k = 0; k < 6; k++) {
        double t = 0.0;
        const double* Rk = s.R.data() + (size_t)k * s.n;
        for (int i = 0; i < s.n; i++) t += Rk[i] * f[i];
        gc[cl.nlc + s.roff + k] = -t;
      }
    }
    cl.Ac.solve(gc.data());
    la[c] = std::move(gc);
  });
  par_for(N, [&](int d) {
    const Sub& s = P.S[d];
    const Cl& cl = P.C[s.cl];
    const std::vector<double>& g = la[s.cl];
    std::vector<double>& x = xs[d];
    std::copy(fs[d].begin(), fs[d].end(), x.begin());
    s.K.solve(x.data());
    const int nb = (int)s.bc_rows.size();
    const double* al = g.data() + cl.nlc + s.roff;
    for (int i = 0; i < s.n; i++) {
      double t = 0.0;
      for (int k = 0; k < 6; k++) t += s.R[(size_t)k * s.n + i] * al[k];
      const double* kb = s.KBcT.data() + (size_t)i * nb;
      for (int q = 0; q < nb; q++) t -= kb[q] * g[s.bc_rows[q]];
      x[i] += t;
    }
  });
}

struct Work {
  std::vector<std::vector<double>> xs, ys;
  explicit Work(const Problem& P) : xs(P.D.nsub), ys(P.D.nsub) {
    for (int d = 0; d < P.D.nsub; d++) xs[d].assign(P.S[d].n, 0.0), ys[d].assign(P.S[d].n, 0.0);
  }
};

// F w = Bf K_c^+ Bf^T w (Cluster.cpp:1114-1188)
void mult_Ff(const Problem& P, Work& W, const double* w, double* Fw) {
  mult_BfT(P, w, W.xs, nullptr);
  mult_Kplus_f(P, W.xs, W.ys);
  mult_Bf(P, W.ys, Fw, nullptr);
}

// M w = W Bf S Bf^T W w (Cluster.cpp:1043-1112); S implicit (Dirichlet_implicit) or K
void precondition(const Problem& P, Work& W, const double* w, double* z) {
  mult_BfT(P, w, W.xs, P.weight.data());
  const int N = P.D.nsub;
  par_for(N, [&](int d) {
    const Sub& s = P.S[d];
    std::vector<double>& y = W.ys[d];
    const std::vector<double>& x = W.xs[d];
    if (P.D.lumped) {
      s.Kfull.apply(x.data(), y.data(), false);
      return;
    }
    const int nb = (int)s.bdofs.size(), nr = (int)s.rdofs.size();
    std::vector<double> xb(nb), yb(nb), t(nr);
    for (int q = 0; q < nb; q++) xb[q] = x[s.bdofs[q]];
    s.Krb.apply(xb.data(), t.data(), false);
    s.Krr.solve(t.data());
    s.Kbb.apply(xb.data(), yb.data(), false);
    std::vector<double> u(nb);
    s.Kbr.apply(t.data(), u.data(), false);
    std::fill(y.begin(), y.end(), 0.0);
    for (int q = 0; q < nb; q++) y[s.bdofs[q]] = yb[q] - u[q];
  });
  mult_Bf(P, W.ys, z, P.weight.data());
}

// P x = x - G (G^T G)^-1 G^T x; alpha = -(G^T G)^-1 G^T x (Cluster.cpp:1314-1329)
void project(const Problem& P, const double* x, double* Px, double* alpha) {
  const int nc = P.ncol;
  const int64_t n = P.nlam;
  std::vector<double> part((size_t)NCHUNK * nc, 0.0);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < NCHUNK; c++) {
    double* pc = part.data() + (size_t)c * nc;
    const int64_t r0 = n * c / NCHUNK, r1 = n * (c + 1) / NCHUNK;
    for (int64_t r = r0; r < r1; r++)
      for (int64_t e = P.rptr[r]; e < P.rptr[r + 1]; e++) {
        const int cl = P.S[P.rent[e].d].cl;
        const double* g = P.Grow.data() + (size_t)e * 6;
        for (int k = 0; k < 6; k++) pc[6 * cl + k] += g[k] * x[r];
      }
  }
  std::vector<double> gt(nc, 0.0), a(nc, 0.0);
  for (int c = 0; c < NCHUNK; c++)
    for (int k = 0; k < nc; k++) gt[k] += part[(size_t)c * nc + k];
  for (int i = 0; i < nc; i++)
    for (int k = 0; k < nc; k++) a[i] += P.invGtG[(size_t)i * nc + k] * gt[k];
  if (alpha)
    for (int i = 0; i < nc; i++) alpha[i] = -a[i];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    double s = 0.0;
    for (int64_t e = P.rptr[r]; e < P.rptr[r + 1]; e++) {
      const int cl = P.S[P.rent[e].d].cl;
      const double* g = P.Grow.data() + (size_t)e * 6;
      for (int k = 0; k < 6; k++) s += g[k] * a[6 * cl + k];
    }
    Px[r] = x[r] - s;
  }
}

void setup(Problem& P, htc_times* tm) {
  const htc_desc& D = P.D;
  const int N = D.nsub, NC = D.ncl;
  P.nlam = D.nlam;
  P.S.resize(N);
  P.weight.assign(D.weight, D.weight + D.nlam);
  auto t0 = clk::now();
  // ---- assembly of K (band, l2g order) and the body load
  par_for(N, [&](int d) {
    Sub& s = P.S[d];
    s.n = (int)(D.n_off[d + 1] - D.n_off[d]);
    const double* X = D.coords + D.n_off[d];
    int kd = 0;
    for (int64_t e = D.e_off[d]; e < D.e_off[d + 1]; e++) {
      const int32_t* c = D.conn + e * 8;
      int lo = c[0], hi = c[0];
      for (int a = 1; a < 8; a++) lo = std::min(lo, c[a]), hi = std::max(hi, c[a]);
      kd = std::max(kd, 3 * (hi - lo) + 2);
    }
    s.K.n = s.n;
    s.K.kd = kd;
    s.K.ab.assign((size_t)(kd + 1) * s.n, 0.0);
    s.f.assign(s.n, 0.0);
    std::vector<std::vector<std::pair<int, double>>> rowsK(s.n);
    double Ke[576], fe[24], Xe[24];
    for (int64_t e = D.e_off[d]; e < D.e_off[d + 1]; e++) {
      const int32_t* c = D.conn + e * 8;
      for (int a = 0; a < 8; a++)
        for (int q = 0; q < 3; q++) Xe[3 * a + q] = X[3 * c[a] + q];
      hex8(Xe, D.Eel[e], D.nu, D.acc, Ke, fe);
      for (int p = 0; p < 24; p++) {
        const int gp = 3 * c[p / 3] + p % 3;
        s.f[gp] += fe[p];
        for (int q = 0; q < 24; q++) {
          const int gq = 3 * c[q / 3] + q % 3;
          if (gp >= gq) s.K.at(gp, gq) += Ke[p * 24 + q];
        }
      }
    }
    // full CSR of K from the band (lumped preconditioner and the Dirichlet blocks)
    s.Kfull.rows = s.n;
    s.Kfull.ptr.assign(s.n + 1, 0);
    for (int i = 0; i < s.n; i++) {
      const int j0 = std::max(0, i - kd), j1 = std::min(s.n - 1, i + kd);
      for (int j = j0; j <= j1; j++) {
        const double v = j <= i ? s.K.at(i, j) : s.K.at(j, i);
        if (v != 0.0) s.Kfull.col.push_back(j), s.Kfull.val.push_back(v);
      }
      s.Kfull.ptr[i + 1] = (int)s.Kfull.col.size();
    }
  });
  tm->assemble = secs(t0);
  // ---- analytic kernel (Data.cpp:61-109), fixing DOFs, K_reg factor
  auto t1 = clk::now();
  par_for(N, [&](int d) {
    Sub& s = P.S[d];
    const double* X = D.coords + D.n_off[d];
    s.R.assign((size_t)s.n * 6, 0.0);
    const int nn = s.n / 3;
    for (int a = 0; a < nn; a++) {
      const double x = X[3 * a], y = X[3 * a + 1], z = X[3 * a + 2];
      double* R = s.R.data();
      R[3 * a + 0] = 1;
      R[3 * a + 1 + s.n] = 1;
      R[3 * a + 2 + 2 * s.n] = 1;
      R[3 * a + 0 + 3 * s.n] = -y, R[3 * a + 1 + 3 * s.n] = x;
      R[3 * a + 0 + 4 * s.n] = -z, R[3 * a + 2 + 4 * s.n] = x;
      R[3 * a + 1 + 5 * s.n] = -z, R[3 * a + 2 + 5 * s.n] = y;
    }
    for (int k = 0; k < 6; k++) {
      double nrm = 0.0;
      for (int i = 0; i < s.n; i++) nrm += s.R[(size_t)k * s.n + i] * s.R[(size_t)k * s.n + i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < s.n; i++) s.R[(size_t)k * s.n + i] /= nrm;
    }
    s.fix = null_pivots(s.R, s.n, 6);
    for (int p : s.fix) s.K.at(p, p) *= 2.0;
    s.K.factor();
  });
  tm->kernel_factor = secs(t1);
  // ---- Dirichlet_implicit: K_rr factor and the K_rb, K_br, K_bb blocks
  auto t2 = clk::now();
  par_for(N, [&](int d) {
    Sub& s = P.S[d];
    std::vector<char> isb(s.n, 0);
    for (int64_t p = D.bf_off[d]; p < D.bf_off[d + 1]; p++) isb[D.bf_col[p]] = 1;
    std::vector<int> pos(s.n);
    for (int i = 0; i < s.n; i++) {
      if (isb[i])
        pos[i] = (int)s.bdofs.size(), s.bdofs.push_back(i);
      else
        pos[i] = (int)s.rdofs.size(), s.rdofs.push_back(i);
    }
    if (D.lumped) return;
    const int nr = (int)s.rdofs.size(), nb = (int)s.bdofs.size();
    int kd = 0;
    for (int a = 0; a < nr; a++) {
      const int i = s.rdofs[a];
      for (int p = s.Kfull.ptr[i]; p < s.Kfull.ptr[i + 1]; p++) {
        const int j = s.Kfull.col[p];
        if (!isb[j] && j < i) kd = std::max(kd, a - pos[j]);
      }
    }
    s.Krr.n = nr;
    s.Krr.kd = kd;
    s.Krr.ab.assign((size_t)(kd + 1) * std::max(nr, 1), 0.0);
    Csr* blocks[3] = {&s.Krb, &s.Kbr, &s.Kbb};
    for (Csr* c : blocks) c->ptr.assign(1, 0);
    s.Krb.rows = nr, s.Kbr.rows = nb, s.Kbb.rows = nb;
    for (int a = 0; a < nr; a++) {
      const int i = s.rdofs[a];
      for (int p = s.Kfull.ptr[i]; p < s.Kfull.ptr[i + 1]; p++) {
        const int j = s.Kfull.col[p];
        if (isb[j])
          s.Krb.col.push_back(pos[j]), s.Krb.val.push_back(s.Kfull.val[p]);
        else if (j <= i)
          s.Krr.at(a, pos[j]) = s.Kfull.val[p];
      }
      s.Krb.ptr.push_back((int)s.Krb.col.size());
    }
    for (int q = 0; q < nb; q++) {
      const int i = s.bdofs[q];
      for (int p = s.Kfull.ptr[i]; p < s.Kfull.ptr[i + 1]; p++) {
        const int j = s.Kfull.col[p];
        if (isb[j])
          s.Kbb.col.push_back(pos[j]), s.Kbb.val.push_back(s.Kfull.val[p]);
        else
          s.Kbr.col.push_back(pos[j]), s.Kbr.val.push_back(s.Kfull.val[p]);
      }
      s.Kbb.ptr.push_back((int)s.Kbb.col.size());
      s.Kbr.ptr.push_back((int)s.Kbr.col.size());
    }
    if (nr) s.Krr.factor();
  });
  tm->dirichlet = secs(t2);
  // ---- cluster objects (Cluster.cpp:151-206) and the cluster matrices
  auto t3 = clk::now();
  P.C.resize(NC);
  for (int c = 0; c < NC; c++) {
    Cl& cl = P.C[c];
    cl.d0 = D.cl_d0[c], cl.d1 = D.cl_d1[c], cl.nlc = D.cl_nlc[c];
    cl.nr = 6 * (cl.d1 - cl.d0);
    for (int d = cl.d0; d < cl.d1; d++) P.S[d].cl = c, P.S[d].roff = 6 * (d - cl.d0);
  }
  par_for(N, [&](int d) {
    Sub& s = P.S[d];
    for (int64_t p = D.bc_off[d]; p < D.bc_off[d + 1]; p++) s.bc_rows.push_back(D.bc_row[p]);
    std::sort(s.bc_rows.begin(), s.bc_rows.end());
    s.bc_rows.erase(std::unique(s.bc_rows.begin(), s.bc_rows.end()), s.bc_rows.end());
    const int nb = (int)s.bc_rows.size();
    s.BcT.assign((size_t)s.n * nb, 0.0);
    for (int64_t p = D.bc_off[d]; p < D.bc_off[d + 1]; p++) {
      const int q = (int)(std::lower_bound(s.bc_rows.begin(), s.bc_rows.end(), D.bc_row[p]) - s.bc_rows.begin());
      s.BcT[(size_t)D.bc_col[p] * nb + q] += D.bc_val[p];
    }
    s.KBcT = s.BcT;
    if (nb) s.K.solve_rows(s.KBcT.data(), nb);
    s.Gc.assign((size_t)nb * 6, 0.0);
    for (int q = 0; q < nb; q++)
      for (int k = 0; k < 6; k++) {
        double t = 0.0;
        for (int i = 0; i < s.n; i++) t += s.BcT[(size_t)i * nb + q] * s.R[(size_t)k * s.n + i];
        s.Gc[(size_t)q * 6 + k] = -t;
      }
  });
  par_for(NC, [&](int c) {
    Cl& cl = P.C[c];
    // H: the global rigid motion k in every subdomain's R basis (R_d column k is the
    // restriction of motion k over its norm), orthonormalised
    cl.H.assign((size_t)cl.nr * 6, 0.0);
    for (int d = cl.d0; d < cl.d1; d++) {
      const Sub& s = P.S[d];
      const double* X = D.coords + D.n_off[d];
      const int nn = s.n / 3;
      double nrm[6] = {};
      for (int a = 0; a < nn; a++) {
        const double x = X[3 * a], y = X[3 * a + 1], z = X[3 * a + 2];
        nrm[0] += 1, nrm[1] += 1, nrm[2] += 1;
        nrm[3] += x * x + y * y, nrm[4] += x * x + z * z, nrm[5] += y * y + z * z;
      }
      for (int k = 0; k < 6; k++) cl.H[(size_t)(s.roff + k) * 6 + k] = std::sqrt(nrm[k]);
    }
    for (int k = 0; k < 6; k++) {  // modified Gram-Schmidt over the 6 columns
      for (int j = 0; j < k; j++) {
        double t = 0.0;
        for (int i = 0; i < cl.nr; i++) t += cl.H[(size_t)i * 6 + j] * cl.H[(size_t)i * 6 + k];
        for (int i = 0; i < cl.nr; i++) cl.H[(size_t)i * 6 + k] -= t * cl.H[(size_t)i * 6 + j];
      }
      double t = 0.0;
      for (int i = 0; i < cl.nr; i++) t += cl.H[(size_t)i * 6 + k] * cl.H[(size_t)i * 6 + k];
      t = 1.0 / std::sqrt(t);
      for (int i = 0; i < cl.nr; i++) cl.H[(size_t)i * 6 + k] *= t;
    }
    // Ac = [Fc Gc 0; Gc^T 0 H; 0 H^T 0] (Cluster.cpp:592-660)
    const int nlc = cl.nlc, nr = cl.nr;
    cl.na = nlc + nr + 6;
    cl.Ac.n = cl.na;
    cl.Ac.a.assign((size_t)cl.na * cl.na, 0.0);
    auto A = [&](int i, int j) -> double& { return cl.Ac.a[(size_t)j * cl.na + i]; };
    for (int d = cl.d0; d < cl.d1; d++) {
      const Sub& s = P.S[d];
      const int nb = (int)s.bc_rows.size();
      for (int q1 = 0; q1 < nb; q1++)
        for (int q2 = 0; q2 < nb; q2++) {
          double t = 0.0;
          for (int i = 0; i < s.n; i++) t += s.BcT[(size_t)i * nb + q1] * s.KBcT[(size_t)i * nb + q2];
          A(s.bc_rows[q1], s.bc_rows[q2]) += t;
        }
      for (int q = 0; q < nb; q++)
        for (int k = 0; k < 6; k++) {
          A(s.bc_rows[q], nlc + s.roff + k) += s.Gc[(size_t)q * 6 + k];
          A(nlc + s.roff + k, s.bc_rows[q]) += s.Gc[(size_t)q * 6 + k];
        }
    }
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < 6; k++) A(nlc + i, nlc + nr + k) = A(nlc + nr + k, nlc + i) = cl.H[(size_t)i * 6 + k];
    cl.Ac.factor();
  });
  tm->cluster = secs(t3);
  // ---- FETI coarse: Rf = R H, Gf = Bf Rf, (Gf^T Gf)^-1
  auto t4 = clk::now();
  P.ncol = 6 * NC;
  par_for(N, [&](int d) {
    Sub& s = P.S[d];
    const Cl& cl = P.C[s.cl];
    s.RH.assign((size_t)s.n * 6, 0.0);
    for (int k = 0; k < 6; k++)
      for (int j = 0; j < 6; j++) {
        const double h = cl.H[(size_t)(s.roff + j) * 6 + k];
        for (int i = 0; i < s.n; i++) s.RH[(size_t)k * s.n + i] += s.R[(size_t)j * s.n + i] * h;
      }
  });
  // dual rows -> entries in subdomain order
  P.rptr.assign(D.nlam + 1, 0);
  for (int d = 0; d < N; d++)
    for (int64_t p = D.bf_off[d]; p < D.bf_off[d + 1]; p++) P.rptr[D.bf_row[p] + 1]++;
  for (int64_t r = 0; r < D.nlam; r++) P.rptr[r + 1] += P.rptr[r];
  P.rent.resize(P.rptr[D.nlam]);
  P.Grow.resize(P.rent.size() * 6);
  {
    std::vector<int64_t> fill(P.rptr.begin(), P.rptr.end() - 1);
    for (int d = 0; d < N; d++)
      for (int64_t p = D.bf_off[d]; p < D.bf_off[d + 1]; p++) {
        const int64_t e = fill[D.bf_row[p]]++;
        P.rent[e] = {d, D.bf_col[p], D.bf_val[p]};
        for (int k = 0; k < 6; k++)
          P.Grow[(size_t)e * 6 + k] = D.bf_val[p] * P.S[d].RH[(size_t)k * P.S[d].n + D.bf_col[p]];
      }
  }
  const int nc = P.ncol;
  std::vector<double> GtG((size_t)nc * nc, 0.0);
  {
    std::vector<double> part((size_t)NCHUNK * nc * nc, 0.0);
    const int64_t n = D.nlam;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; c++) {
      double* pc = part.data() + (size_t)c * nc * nc;
      const int64_t r0 = n * c / NCHUNK, r1 = n * (c + 1) / NCHUNK;
      for (int64_t r = r0; r < r1; r++) {
        // the row of Gf: 6 values per cluster its entries belong to (at most 2)
        int ncl = 0, cls[4];
        double v[4][6];
        for (int64_t e = P.rptr[r]; e < P.rptr[r + 1]; e++) {
          const int cl = P.S[P.rent[e].d].cl;
          int q = 0;
          while (q < ncl && cls[q] != cl) q++;
          if (q == ncl) {
            cls[ncl++] = cl;
            for (int k = 0; k < 6; k++) v[q][k] = 0.0;
          }
          for (int k = 0; k < 6; k++) v[q][k] += P.Grow[(size_t)e * 6 + k];
        }
        for (int a = 0; a < ncl; a++)
          for (int b = 0; b < ncl; b++)
            for (int i = 0; i < 6; i++)
              for (int j = 0; j < 6; j++) pc[(size_t)(6 * cls[a] + i) * nc + 6 * cls[b] + j] += v[a][i] * v[b][j];
      }
    }
    for (int c = 0; c < NCHUNK; c++)
      for (size_t q = 0; q < (size_t)nc * nc; q++) GtG[q] += part[(size_t)c * nc * nc + q];
  }
  // Cholesky of GtG, inverse by solving against I, symmetrised (Cluster.cpp:1246-1273)
  std::vector<double> L(GtG);
  for (int j = 0; j < nc; j++) {
    double s = L[(size_t)j * nc + j];
    for (int k = 0; k < j; k++) s -= L[(size_t)j * nc + k] * L[(size_t)j * nc + k];
    if (!(s > 0.0)) throw std::runtime_error("Gf^T Gf not positive definite");
    L[(size_t)j * nc + j] = std::sqrt(s);
    for (int i = j + 1; i < nc; i++) {
      double t = L[(size_t)i * nc + j];
      for (int k = 0; k < j; k++) t -= L[(size_t)i * nc + k] * L[(size_t)j * nc + k];
      L[(size_t)i * nc + j] = t / L[(size_t)j * nc + j];
    }
  }
  P.invGtG.assign((size_t)nc * nc, 0.0);
  for (int c = 0; c < nc; c++) {
    std::vector<double> x(nc, 0.0);
    x[c] = 1.0;
    for (int i = 0; i < nc; i++) {
      double t = x[i];
      for (int k = 0; k < i; k++) t -= L[(size_t)i * nc + k] * x[k];
      x[i] = t / L[(size_t)i * nc + i];
    }
    for (int i = nc - 1; i >= 0; i--) {
      double t = x[i];
      for (int k = i + 1; k < nc; k++) t -= L[(size_t)k * nc + i] * x[k];
      x[i] = t / L[(size_t)i * nc + i];
    }
    for (int i = 0; i < nc; i++) P.invGtG[(size_t)i * nc + c] = x[i];
  }
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < i; j++) {
      const double v = 0.5 * (P.invGtG[(size_t)i * nc + j] + P.invGtG[(size_t)j * nc + i]);
      P.invGtG[(size_t)i * nc + j] = P.invGtG[(size_t)j * nc + i] = v;
    }
  tm->coarse = secs(t4);
  tm->total = secs(t0);
}

// Cluster::pcpg (Cluster.cpp:1363-1458) and the recovery (Cluster.cpp:1648-1678)
void solve(const Problem& P, const std::vector<std::vector<double>>& f, double eps, int max_iter, double* u,
           int* iters, double* resid, htc_solve_times* tm) {
  const int64_t n = P.nlam;
  const int N = P.D.nsub, nc = P.ncol;
  Work W(P);
  auto t0 = clk::now();
  std::vector<std::vector<double>> xx(N);
  for (int d = 0; d < N; d++) xx[d].assign(P.S[d].n, 0.0);
  mult_Kplus_f(P, f, xx);
  std::vector<double> drhs(n), lam(n), g(n), Pg(n), z(n), Pz(n), w(n), Fw(n), alpha(nc), e(nc, 0.0);
  mult_Bf(P, xx, drhs.data(), nullptr);
  for (int d = 0; d < N; d++) {  // e = Rf^T f
    const Sub& s = P.S[d];
    for (int k = 0; k < 6; k++) {
      double t = 0.0;
      for (int i = 0; i < s.n; i++) t += s.RH[(size_t)k * s.n + i] * f[d][i];
      e[6 * s.cl + k] += t;
    }
  }
  std::vector<double> a(nc, 0.0);
  for (int i = 0; i < nc; i++)
    for (int k = 0; k < nc; k++) a[i] += P.invGtG[(size_t)i * nc + k] * e[k];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {  // lam0 = Gf (Gf^T Gf)^-1 e
    double s = 0.0;
    for (int64_t q = P.rptr[r]; q < P.rptr[r + 1]; q++) {
      const int cl = P.S[P.rent[q].d].cl;
      for (int k = 0; k < 6; k++) s += P.Grow[(size_t)q * 6 + k] * a[6 * cl + k];
    }
    lam[r] = s;
  }
  mult_Ff(P, W, lam.data(), g.data());
  for (int64_t r = 0; r < n; r++) g[r] -= drhs[r];
  project(P, g.data(), Pg.data(), alpha.data());
  precondition(P, W, Pg.data(), z.data());
  project(P, z.data(), Pz.data(), nullptr);
  double gPz = dot(P, g.data(), Pz.data());
  const double norm0 = std::sqrt(gPz);
  w = Pz;
  tm->init = secs(t0);
  auto t1 = clk::now();
  int it = 0;
  for (it = 0; it <= max_iter; it++) {
    if (it == max_iter || std::sqrt(gPz) < eps * norm0) break;
    mult_Ff(P, W, w.data(), Fw.data());
    const double rho = -gPz / dot(P, w.data(), Fw.data());
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) lam[r] += rho * w[r], g[r] += rho * Fw[r];
    project(P, g.data(), Pg.data(), alpha.data());
    precondition(P, W, Pg.data(), z.data());
    project(P, z.data(), Pz.data(), nullptr);
    const double prev = gPz;
    gPz = dot(P, g.data(), Pz.data());
    const double gamma = gPz / prev;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; r++) w[r] = Pz[r] + gamma * w[r];
  }
  tm->pcpg = secs(t1);
  auto t2 = clk::now();
  mult_BfT(P, lam.data(), W.xs, nullptr);
  for (int d = 0; d < N; d++)
    for (int i = 0; i < P.S[d].n; i++) W.xs[d][i] = f[d][i] - W.xs[d][i];
  mult_Kplus_f(P, W.xs, xx);
  for (int d = 0; d < N; d++) {  // later subdomains overwrite shared DOFs
    const Sub& s = P.S[d];
    const int64_t* l2g = P.D.l2g + P.D.n_off[d];
    for (int i = 0; i < s.n; i++) {
      double t = xx[d][i];
      for (int k = 0; k < 6; k++) t -= s.RH[(size_t)k * s.n + i] * alpha[6 * s.cl + k];
      u[l2g[i]] = t;
    }
  }
  tm->recover = secs(t2);
  tm->total = secs(t0);
  *iters = it;
  *resid = norm0 > 0 ? std::sqrt(std::max(gPz, 0.0)) / norm0 : 0.0;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* htc_last_error() { return g_err.c_str(); }

// set-up with `threads` OpenMP threads; returns a handle or NULL (htc_last_error)
void* htc_setup(const htc_desc* desc, int threads, htc_times* tm) {
  try {
    omp_set_num_threads(threads);
    std::unique_ptr<Problem> P(new Problem());
    P->D = *desc;
    setup(*P, tm);
    return P.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

// one load case: f packed per subdomain in l2g order (NULL: the body load);
// u has desc->ndof entries
int htc_solve(void* h, const double* f, double eps, int max_iter, int threads, double* u, int* iters,
              double* resid, htc_solve_times* tm) {
  try {
    omp_set_num_threads(threads);
    const Problem& P = *(const Problem*)h;
    std::vector<std::vector<double>> fs(P.D.nsub);
    for (int d = 0; d < P.D.nsub; d++) {
      if (f)
        fs[d].assign(f + P.D.n_off[d], f + P.D.n_off[d + 1]);
      else
        fs[d] = P.S[d].f;
    }
    solve(P, fs, eps, max_iter, u, iters, resid, tm);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// the body load (packed, l2g order)
void htc_body_load(void* h, double* f) {
  const Problem& P = *(const Problem*)h;
  for (int d = 0; d < P.D.nsub; d++) std::copy(P.S[d].f.begin(), P.S[d].f.end(), f + P.D.n_off[d]);
}

// the fixing DOFs of subdomain d (6)
void htc_fix(void* h, int d, int* out) {
  const Problem& P = *(const Problem*)h;
  for (int k = 0; k < 6; k++) out[k] = P.S[d].fix[k];
}

void htc_free(void* h) { delete (Problem*)h; }

}  // extern "C"
