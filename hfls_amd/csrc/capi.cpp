// C ABI (include/hfls_gpu.h) over the host problem and the device engine.
#include <cstring>
#include <memory>
#include <exception>
#include <stdexcept>
#include <string>

#include "../../include/hfls_gpu.h"
#include "engine.hpp"
#include "problem.hpp"
#include "serve.hpp"
#include "state.hpp"
#include "vtk.hpp"

struct hfls_problem_s {
  hfls::Problem* p;
};
struct hfls_mesh_s {
  hfls::Engine* e;
};
struct hfls_server_s {
  hfls::Server* s;
};

static thread_local std::string g_err;

extern "C" const char* hfls_last_error(void) { return g_err.c_str(); }

template <class F>
static int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -2;
  }
}

// the object behind a handle; a NULL handle is an error, not a crash
static hfls::Problem* prob(hfls_problem h) {
  if (!h || !h->p) throw std::runtime_error("null problem handle");
  return h->p;
}
static hfls::Engine* eng(hfls_mesh m) {
  if (!m || !m->e) throw std::runtime_error("null mesh handle");
  return m->e;
}
static hfls::Server* srv(hfls_server s) {
  if (!s || !s->s) throw std::runtime_error("null server handle");
  return s->s;
}
template <class T>
static T* need(T* p, const char* what) {
  if (!p) throw std::runtime_error(std::string("null argument: ") + what);
  return p;
}

static hfls::Options make_options(int n, const char* const* k, const char* const* v) {
  hfls::Options o = hfls::default_options();
  if (n > 0 && (!k || !v)) throw std::runtime_error("null option arrays");
  for (int i = 0; i < n; i++) {
    if (!k[i] || !v[i]) throw std::runtime_error("null option key or value");
    if (!o.count(k[i])) throw std::runtime_error(std::string("unknown option ") + k[i]);
    o[k[i]] = v[i];
  }
  return o;
}

extern "C" int hfls_problem_create(int nopt, const char* const* keys, const char* const* vals, hfls_problem* out) {
  return guard([&] {
    need(out, "out");
    std::unique_ptr<hfls::Problem> p(new hfls::Problem(make_options(nopt, keys, vals)));
    auto* h = new hfls_problem_s;  // after the problem is built: nothing to leak if it throws
    h->p = p.release();
    *out = h;
  });
}

extern "C" int hfls_problem_destroy(hfls_problem p) {
  return guard([&] {
    if (p) {
      delete p->p;
      delete p;
    }
  });
}

extern "C" int hfls_problem_sizes(hfls_problem h, int64_t* s) {
  return guard([&] {
    need(s, "sizes");
    const hfls::Problem& p = *prob(h);
    s[0] = p.nsub; s[1] = p.nlam; s[2] = p.n_glue; s[3] = p.nPoints; s[4] = p.ncl; s[5] = p.nsub_cl;
    s[6] = (int64_t)p.dirichlet_dofs.size(); s[7] = (int64_t)p.corner_dofs.size();
  });
}

static const hfls::Subdomain& sub(hfls_problem h, int d) {
  if (d < 0 || d >= prob(h)->nsub) throw std::runtime_error("subdomain index out of range");
  return prob(h)->subs[d];
}

extern "C" int hfls_problem_sub_info(hfls_problem h, int d, int* info) {
  return guard([&] {
    need(info, "info");
    const hfls::Subdomain& s = sub(h, d);
    info[0] = s.n; info[1] = s.n_r; info[2] = s.n_s; info[3] = s.n_bs; info[4] = s.n_ss; info[5] = s.cluster;
    info[6] = (int)s.bf_row.size(); info[7] = 0;
  });
}

extern "C" int hfls_problem_sub_bf(hfls_problem h, int d, int* rows, int* dofs, int* signs) {
  return guard([&] {
    need(rows, "rows"); need(dofs, "dofs"); need(signs, "signs");
    const hfls::Subdomain& s = sub(h, d);
    for (size_t i = 0; i < s.bf_row.size(); i++) {
      rows[i] = s.bf_row[i];
      dofs[i] = s.bf_dof[i];
      signs[i] = s.bf_sign[i];
    }
  });
}

extern "C" int hfls_problem_sub_bdofs(hfls_problem h, int d, int* b) {
  return guard([&] {
    need(b, "bdofs");
    const hfls::Subdomain& s = sub(h, d);
    std::memcpy(b, s.bdofs.data(), s.bdofs.size() * sizeof(int));
  });
}

extern "C" int hfls_problem_sub_sset(hfls_problem h, int d, int* b) {
  return guard([&] {
    need(b, "sset");
    const hfls::Subdomain& s = sub(h, d);
    std::memcpy(b, s.sset.data(), s.sset.size() * sizeof(int));
  });
}

extern "C" int hfls_problem_sub_l2g(hfls_problem h, int d, int64_t* l2g) {
  return guard([&] {
    need(l2g, "l2g");
    const hfls::Problem& p = *prob(h);
    const hfls::Subdomain& s = sub(h, d);
    const int64_t nx1 = p.NE[0] + 1, nxy1 = nx1 * (p.NE[1] + 1);
    for (int l = 0; l < s.n; l++) {
      int ln = l / 3, c = l % 3;
      int64_t ii = s.e0[0] + ln % s.nn[0], jj = s.e0[1] + (ln / s.nn[0]) % s.nn[1],
              kk = s.e0[2] + ln / (s.nn[0] * s.nn[1]);
      l2g[l] = 3 * (ii + jj * nx1 + kk * nxy1) + c;
    }
  });
}

extern "C" int hfls_problem_sub_young(hfls_problem h, int d, double* E) {
  return guard([&] { *need(E, "E") = sub(h, d).E; });
}

extern "C" int hfls_problem_write_vtk(hfls_problem h, const double* u, const char* folder, int iter) {
  return guard([&] {
    if (!u || !folder) throw std::runtime_error("null argument");
    hfls::write_vtk(*prob(h), u, folder, iter);
  });
}

extern "C" int hfls_problem_rows(hfls_problem h, int* a, int* b, int64_t* dof, double* w) {
  return guard([&] {
    const hfls::Problem& p = *prob(h);
    for (int64_t r = 0; r < p.nlam; r++) {
      if (a) a[r] = p.row_sub_a[r];
      if (b) b[r] = p.row_sub_b[r];
      if (dof) dof[r] = p.row_dof[r];
      if (w) w[r] = p.weight[r];
    }
  });
}

extern "C" int hfls_problem_dirichlet(hfls_problem h, int64_t* dofs) {
  return guard([&] {
    need(dofs, "dofs");
    const auto& v = prob(h)->dirichlet_dofs;
    std::memcpy(dofs, v.data(), v.size() * sizeof(int64_t));
  });
}

extern "C" int hfls_problem_corners(hfls_problem h, int64_t* dofs) {
  return guard([&] {
    need(dofs, "dofs");
    const auto& v = prob(h)->corner_dofs;
    std::memcpy(dofs, v.data(), v.size() * sizeof(int64_t));
  });
}

// ------------------------------------------------------------------ device mesh
extern "C" int hfls_mesh_create(int device, int nopt, const char* const* keys, const char* const* vals,
                                hfls_mesh* out) {
  return guard([&] {
    need(out, "out");
    hfls::Options o = make_options(nopt, keys, vals);
    std::unique_ptr<hfls::Engine> e(new hfls::Engine(o, device));
    e->setup();
    auto* h = new hfls_mesh_s;
    h->e = e.release();
    *out = h;
  });
}

extern "C" int hfls_mesh_save(hfls_mesh m, const char* path) {
  return guard([&] { hfls::save_mesh_file(*eng(m), need(path, "path")); });
}

extern "C" int hfls_mesh_restore(int device, const char* path, hfls_mesh* out, double* restore_ms) {
  return guard([&] {
    need(out, "out");
    std::unique_ptr<hfls::Engine> e = hfls::load_mesh_file(need(path, "path"), device);
    if (restore_ms) *restore_ms = e->setup_info().restore_ms;
    auto* h = new hfls_mesh_s;
    h->e = e.release();
    *out = h;
  });
}

extern "C" int hfls_mesh_destroy(hfls_mesh m) {
  return guard([&] {
    if (m) {
      delete m->e;
      delete m;
    }
  });
}

extern "C" int hfls_mesh_sizes(hfls_mesh m, int64_t* s) {
  return guard([&] {
    need(s, "sizes");
    const hfls::Engine& e = *eng(m);
    s[0] = e.problem().nsub; s[1] = e.nlam(); s[2] = e.ncol_G(); s[3] = e.packed_n();
    s[4] = e.problem().nPoints; s[5] = (int64_t)e.setup_info().device_bytes;
  });
}

extern "C" int hfls_mesh_setup_times(hfls_mesh m, double* t) {
  return guard([&] {
    need(t, "times");
    const hfls::SetupInfo& i = eng(m)->setup_info();
    t[0] = i.setup_ms;
    t[1] = i.factor_ms;
    t[2] = i.meta_ms;
    t[3] = i.assemble_ms;
    t[4] = i.coarse_ms;
  });
}

extern "C" int hfls_mesh_write_vtk(hfls_mesh m, const double* u, const char* folder, int iter) {
  return guard([&] {
    if (!u || !folder) throw std::runtime_error("null argument");
    hfls::write_vtk(eng(m)->problem(), u, folder, iter);
  });
}

extern "C" int hfls_mesh_kernel(hfls_mesh m, double* R, int* fix, int* defect) {
  return guard([&] {
    std::vector<double> r;
    std::vector<int> f, d;
    if (R) {
      eng(m)->get_R(r);
      std::memcpy(R, r.data(), r.size() * sizeof(double));
    }
    if (fix) {
      eng(m)->get_fix(f);
      std::memcpy(fix, f.data(), f.size() * sizeof(int));
    }
    if (defect) {
      eng(m)->get_defect(d);
      std::memcpy(defect, d.data(), d.size() * sizeof(int));
    }
  });
}

extern "C" int hfls_mesh_load(hfls_mesh m, double* f) {
  return guard([&] {
    need(f, "f");
    std::vector<double> v;
    eng(m)->get_load(v);
    std::memcpy(f, v.data(), v.size() * sizeof(double));
  });
}

extern "C" int hfls_apply_F(hfls_mesh m, const double* w, double* Fw) {
  return guard([&] { eng(m)->apply_F(need(w, "w"), need(Fw, "Fw")); });
}

extern "C" int hfls_project(hfls_mesh m, const double* x, double* Px, double* alpha) {
  return guard([&] { eng(m)->project(need(x, "x"), need(Px, "Px"), alpha); });
}

extern "C" int hfls_precondition(hfls_mesh m, const double* x, double* Mx) {
  return guard([&] { eng(m)->precondition(need(x, "x"), need(Mx, "Mx")); });
}

extern "C" int hfls_kplus(hfls_mesh m, const double* f, double* x) {
  return guard([&] { eng(m)->kplus_local(need(f, "f"), need(x, "x")); });
}

extern "C" int hfls_mult_Gf(hfls_mesh m, const double* alpha, double* lam) {
  return guard([&] { eng(m)->mult_Gf(need(alpha, "alpha"), need(lam, "lambda")); });
}

extern "C" int hfls_mult_GfT(hfls_mesh m, const double* lam, double* e) {
  return guard([&] { eng(m)->mult_GfT(need(lam, "lambda"), need(e, "e")); });
}

extern "C" int hfls_mult_RfT(hfls_mesh m, const double* f, double* e) {
  return guard([&] { eng(m)->mult_RfT(need(f, "f_local"), need(e, "e")); });
}

extern "C" int hfls_recover(hfls_mesh m, const double* f, const double* lam, const double* alpha, double* x_local,
                            double* u) {
  return guard([&] { eng(m)->recover(need(f, "f_local"), need(lam, "lambda"), need(alpha, "alpha"), x_local, u); });
}

extern "C" int hfls_project_noalloc(hfls_mesh m, const double* x_dev, double* Px_dev, double* alpha_dev,
                                    void* stream) {
  return guard([&] {
    eng(m)->project_noalloc(need(x_dev, "x"), need(Px_dev, "Px"), alpha_dev, static_cast<hipStream_t>(stream));
  });
}

extern "C" int hfls_mesh_rf(hfls_mesh m, double* Rf) {
  return guard([&] {
    std::vector<double> r;
    eng(m)->get_Rf(r);
    std::memcpy(need(Rf, "Rf"), r.data(), r.size() * sizeof(double));
  });
}

extern "C" int hfls_mesh_counts(hfls_mesh m, int64_t* c) {
  return guard([&] {
    need(c, "counts");
    const hfls::Engine& e = *eng(m);
    c[0] = e.problem().nsub;
    c[1] = e.packed_n();
    c[2] = e.nlam_c();
    c[3] = e.nlam();
    c[4] = e.ncol_G();
    c[5] = 6 * (int64_t)e.problem().nsub;
  });
}

extern "C" int hfls_solve(hfls_mesh m, double eps, int max_iter, const double* f, double* u, double* lam,
                          hfls_stats* stats) {
  return guard([&] {
    hfls::SolveInfo info;
    eng(m)->solve(eps, max_iter, f, u, lam, &info);
    if (stats) {
      stats->iterations = info.iterations;
      stats->converged = info.converged;
      stats->residual = info.residual;
      stats->solve_ms = info.solve_ms;
      stats->pcpg_ms = info.pcpg_ms;
      stats->pg_rel = info.pg_rel;
    }
  });
}

extern "C" int hfls_time_stream(hfls_mesh m, int which, int reps, double* ms, double* bytes) {
  return guard([&] { *need(ms, "ms") = eng(m)->time_stream(which, reps, bytes); });
}

extern "C" int hfls_solve_block(hfls_mesh m, double eps, int max_iter, int k, int precision, const double* f,
                                double* u, double* lam, hfls_case_stats* cases, hfls_block_stats* stats) {
  return guard([&] {
    if (k < 1 || k > hfls::KMAX) throw std::runtime_error("k must be in 1..64");
    std::vector<int> it(k), stt(k);
    std::vector<double> res(k);
    hfls::BlockInfo info;
    eng(m)->solve_block(eps, max_iter, k, precision, f, u, lam, it.data(), stt.data(), res.data(), &info);
    if (cases)
      for (int j = 0; j < k; j++) {
        cases[j].iterations = it[j];
        cases[j].status = stt[j];
        cases[j].residual = res[j];
      }
    if (stats) {
      stats->solve_ms = info.solve_ms;
      stats->init_ms = info.init_ms;
      stats->pcpg_ms = info.pcpg_ms;
      stats->recover_ms = info.recover_ms;
      stats->sweeps = info.sweeps;
      stats->cert_steps = info.cert_steps;
    }
  });
}

// page-locked host memory for load and result arrays: device copies from and to it
// are DMA transfers at the link rate, not staged through the runtime's buffers
extern "C" int hfls_host_alloc(int64_t bytes, void** out) {
  return guard([&] {
    if (!out) throw std::runtime_error("hfls_host_alloc: null output pointer");
    *out = nullptr;
    if (bytes <= 0) throw std::runtime_error("hfls_host_alloc: bytes must be positive");
    void* p = nullptr;
    const hipError_t e = hipHostMalloc(&p, (size_t)bytes, hipHostMallocDefault);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
    *out = p;
  });
}

extern "C" int hfls_host_free(void* p) {
  return guard([&] {
    if (!p) return;
    const hipError_t e = hipHostFree(p);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipHostFree: ") + hipGetErrorString(e));
  });
}

extern "C" int hfls_solve_blocks(hfls_mesh m, double eps, int max_iter, int nblocks, int k, int precision,
                                 const double* f, double* u, hfls_case_stats* cases, hfls_block_stats* stats) {
  return guard([&] {
    if (k < 1 || k > hfls::KMAX) throw std::runtime_error("k must be in 1..64");
    if (nblocks < 1 || nblocks > 8) throw std::runtime_error("nblocks must be in 1..8");
    if (!f || !u) throw std::runtime_error("null argument");
    const size_t kt = (size_t)nblocks * k;
    std::vector<int> it(kt), stt(kt);
    std::vector<double> res(kt);
    hfls::BlockInfo info;
    eng(m)->solve_blocks(nblocks, eps, max_iter, k, precision, f, u, it.data(), stt.data(), res.data(), &info);
    if (cases)
      for (size_t j = 0; j < kt; j++) {
        cases[j].iterations = it[j];
        cases[j].status = stt[j];
        cases[j].residual = res[j];
      }
    if (stats) {
      stats->solve_ms = info.solve_ms;
      stats->init_ms = info.init_ms;
      stats->pcpg_ms = info.pcpg_ms;
      stats->recover_ms = info.recover_ms;
      stats->sweeps = info.sweeps;
      stats->cert_steps = info.cert_steps;
    }
  });
}

extern "C" int hfls_apply_block(hfls_mesh m, int op, int k, int precision, const double* x, double* y) {
  return guard([&] { eng(m)->apply_block(op, k, precision, need(x, "x"), need(y, "y")); });
}

extern "C" int hfls_problem_traction(hfls_problem p, uint64_t seed, int case0, int k, double* f) {
  return guard([&] {
    if (k < 1 || case0 < 0) throw std::runtime_error("invalid load-case range");
    hfls::traction_loads(*prob(p), seed, case0, k, need(f, "f"));
  });
}

extern "C" int hfls_mesh_traction(hfls_mesh m, uint64_t seed, int case0, int k, double* f) {
  return guard([&] {
    if (k < 1 || case0 < 0) throw std::runtime_error("invalid load-case range");
    hfls::traction_loads(eng(m)->problem(), seed, case0, k, need(f, "f"));
  });
}

// ------------------------------------------------------------------ server
extern "C" int hfls_server_create(int device, double budget_gb, int slots, int precision, double eps, int max_iter,
                                  int chunk, hfls_server* out) {
  return guard([&] {
    need(out, "out");
    if (!(budget_gb > 0)) throw std::runtime_error("HBM budget must be positive");
    std::unique_ptr<hfls::Server> sv(
        new hfls::Server(device, (size_t)(budget_gb * 1e9), slots, precision, eps, max_iter, chunk));
    auto* h = new hfls_server_s;
    h->s = sv.release();
    *out = h;
  });
}

extern "C" int hfls_server_destroy(hfls_server s) {
  return guard([&] {
    if (s) {
      delete s->s;
      delete s;
    }
  });
}

extern "C" int hfls_server_sizes(hfls_server s, int nopt, const char* const* keys, const char* const* vals,
                                 int64_t* npacked, int64_t* nglob, int64_t* bytes) {
  return guard([&] { srv(s)->sizes(make_options(nopt, keys, vals), npacked, nglob, bytes); });
}

extern "C" int hfls_server_submit(hfls_server s, int nopt, const char* const* keys, const char* const* vals, int k,
                                  const double* f, uint64_t seed, int case0, int64_t* id) {
  return guard([&] { *need(id, "id") = srv(s)->submit(make_options(nopt, keys, vals), k, f, seed, case0); });
}

extern "C" int hfls_server_step(hfls_server s, int* busy) {
  return guard([&] {
    const bool b = srv(s)->step();
    if (busy) *busy = b ? 1 : 0;
  });
}

extern "C" int hfls_server_state(hfls_server s, int64_t id, int* state) {
  return guard([&] { *need(state, "state") = srv(s)->state(id); });
}

extern "C" int hfls_server_result(hfls_server s, int64_t id, double* u, hfls_case_stats* cases, double* latency_ms) {
  return guard([&] {
    const int k = srv(s)->request_k(id);
    std::vector<int> it(k), stt(k);
    std::vector<double> res(k);
    double lat[2] = {0, 0};
    auto copy_out = [&] {
      if (cases)
        for (int j = 0; j < k; j++) {
          cases[j].iterations = it[j];
          cases[j].status = stt[j];
          cases[j].residual = res[j];
        }
      if (latency_ms) {
        latency_ms[0] = lat[0];
        latency_ms[1] = lat[1];
      }
    };
    try {
      srv(s)->result(id, u, it.data(), stt.data(), res.data(), lat, nullptr);
    } catch (...) {
      copy_out();  // a failed set-up: statuses 7, then the error
      throw;
    }
    copy_out();
  });
}

extern "C" int hfls_server_cancel(hfls_server s, int64_t id) {
  return guard([&] { srv(s)->cancel(id); });
}

extern "C" int hfls_server_stats(hfls_server s, double* out) {
  return guard([&] {
    need(out, "out");
    const hfls::ServeStats& t = srv(s)->stats();
    out[0] = srv(s)->meshes();
    out[1] = (double)srv(s)->bytes_in_use();
    out[2] = (double)t.setups;
    out[3] = (double)t.evictions;
    out[4] = (double)t.cases_done;
    out[5] = (double)t.requests_done;
    out[6] = (double)t.joins;
    out[7] = t.setup_ms;
    out[8] = (double)t.failed_setups;
    out[9] = (double)t.cancelled;
    out[10] = (double)t.steps;
    out[11] = t.join_ms;
    out[12] = t.sweep_ms;
    out[13] = t.finish_ms;
  });
}

extern "C" int hfls_server_host_cache(hfls_server s, double gb) {
  return guard([&] {
    if (!(gb >= 0.0) || gb > 1e6) throw std::runtime_error("host cache size must be 0..1e6 GB");
    srv(s)->set_host_cache((size_t)(gb * 1e9));
  });
}

extern "C" int hfls_server_cache_stats(hfls_server s, double* out) {
  return guard([&] {
    need(out, "out");
    const hfls::ServeStats& t = srv(s)->stats();
    out[0] = (double)t.host_meshes;
    out[1] = (double)t.host_bytes;
    out[2] = (double)t.saves;
    out[3] = t.save_ms;
    out[4] = (double)t.restores;
    out[5] = t.restore_ms;
    out[6] = (double)t.host_drops;
  });
}

extern "C" int hfls_mesh_configure(hfls_mesh m, const char* key, const char* value) {
  return guard([&] {
    if (!key || !value) throw std::runtime_error("null setting");
    eng(m)->configure(key, value);
  });
}

extern "C" int hfls_mesh_tuning(hfls_mesh m, double* out) {
  return guard([&] {
    need(out, "out");
    out[0] = eng(m)->tuned_rows();
    out[1] = eng(m)->tuned_ms(1);
    out[2] = eng(m)->tuned_ms(2);
    out[3] = eng(m)->tuned_ms(3);
  });
}
