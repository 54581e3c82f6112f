// Device-resident HTFETI mesh: setup (assembly, factorisation, kernel detection,
// explicit K+_bb and S_i, HTFETI cluster inverse, projector) and PCPG solves.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "problem.hpp"

namespace hfls {

// throw on a failed HIP call, naming the call and its place
#define HC(x)                                                                                  \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess)                                                                      \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +    \
                               __FILE__ + ":" + std::to_string(__LINE__) + ": " #x);           \
  } while (0)

// HIP events destroyed when the scope ends, also when a HIP call in it throws
struct Events {
  std::vector<hipEvent_t> e;
  explicit Events(size_t n) : e(n, nullptr) {
    for (auto& x : e) {
      const hipError_t r = hipEventCreate(&x);
      if (r != hipSuccess) {
        x = nullptr;
        release();
        throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(r) + ": hipEventCreate");
      }
    }
  }
  Events(const Events&) = delete;
  Events& operator=(const Events&) = delete;
  ~Events() { release(); }
  void release() {
    for (auto& x : e)
      if (x) {
        (void)hipEventDestroy(x);
        x = nullptr;
      }
  }
  hipEvent_t operator[](size_t i) const { return e[i]; }
};

struct SolveInfo {
  int iterations = 0;
  double residual = 0.0;     // sqrt(g'Pz)/sqrt(g0'Pz0) at exit
  double pg_rel = 0.0;       // |P g| / |P g0| at exit
  double norm0 = 0.0;
  double solve_ms = 0.0;     // device time of the whole load-case solve
  double pcpg_ms = 0.0;      // device time of the iteration loop
  double gemv_ms = 0.0;      // summed device time of the K+_bb stream kernels (F apply)
  int gemv_launches = 0;
  int converged = 0;
};

struct SetupInfo {
  double setup_ms = 0.0;
  double factor_ms = 0.0;    // partial Cholesky, kernel detection, K+_bb (host wall, synchronised)
  double meta_ms = 0.0;      // index structures
  double assemble_ms = 0.0;  // scratch allocation and element assembly
  double coarse_ms = 0.0;    // Bc, cluster and coarse problems
  double restore_ms = 0.0;   // a mesh loaded from a saved state: the load (the others: the original set-up)
  size_t device_bytes = 0;
  int min_defect = 0, max_defect = 0;
};

// Block solve of k <= 64 load cases (one column each).
struct BlockInfo {
  double solve_ms = 0.0, init_ms = 0.0, pcpg_ms = 0.0, recover_ms = 0.0, cert_ms = 0.0;
  int sweeps = 0;        // block iterations executed
  int cert_steps = 0;    // FP64 true-residual evaluations summed over the columns
};

template <class T>
struct DVec {
  T* p = nullptr;
  size_t n = 0;
  DVec() = default;
  DVec(const DVec&) = delete;
  DVec& operator=(const DVec&) = delete;
  ~DVec() { release(); }
  void release();
  void alloc(size_t count);
  void upload(const std::vector<T>& v, hipStream_t st);
  void download(std::vector<T>& v, hipStream_t st) const;
  void swap(DVec& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

// The state of one block solve: the k columns' buffers and scalars, its stream and
// its captured sweep graph.  Several contexts of one mesh run blocks concurrently
// on their own streams and share the mesh's operators (tiles, factors, index
// arrays); a context is brought into the engine with Engine::use_ctx.
struct BlockCtx {
  hipStream_t st = nullptr;
  hipGraphExec_t gx = nullptr;
  int64_t gx_key = -1;
  int k = 0, prec = 64;
  DVec<float> xb32, yb32, sx, p16;
  DVec<double> gtp;
  DVec<double> xb, yb, cr, coef, kpart, v6, gc, la, h, hg, hgf, lam, g, w, Fw, Pg, z, Pz, d, g64, part, gt, Vr, Xr, Tr,
      V, V2, X, U;
  DVec<BScalars> sc;
};

class StateStream;

class Engine {
 public:
  Engine(const Options& o, int device);
  ~Engine();
  void setup();
  // The set-up state (the products of setup(): the factors W = L_rr^-1, the K+_bb and
  // S_i tiles, kernels and fixing DOFs, the gluing and the coarse inverses) through a
  // byte stream: save_state after setup(); load_state on an Engine constructed with the
  // options the state was saved with, in place of setup().  The index structures are
  // rebuilt from the options, every loaded size and index is checked against them, and
  // a loaded mesh computes bit-identical results to the mesh that was saved.
  // copy: the stream of the device-to-host copies of a save (null: the engine's own;
  // the saved state is immutable, so another stream may copy it while the engine runs)
  void save_state(StateStream& s, hipStream_t copy = nullptr);
  void load_state(StateStream& s);
  const Problem& problem() const { return *P; }
  int64_t nlam() const { return P->nlam; }
  int ncol_G() const { return 6 * P->ncl; }
  const SetupInfo& setup_info() const { return sinfo; }

  // Operators on host arrays (FP64).  Parity entry points of the C ABI.
  void apply_F(const double* w, double* Fw);
  void project(const double* x, double* Px, double* alpha);
  void precondition(const double* x, double* Mx);
  void kplus_local(const double* f_local, double* x_local);  // K_c^+ f (packed local vectors)
  // Gf alpha (Cluster.cpp:1275-1283), Gf^T lambda (1285-1293), Rf^T f (1295-1312);
  // coarse vectors have 6 entries per cluster
  void mult_Gf(const double* alpha, double* lam);
  void mult_GfT(const double* lam, double* e);
  void mult_RfT(const double* f_local, double* e);
  // recovery of printVTK (Cluster.cpp:1648-1678) from host lambda and alpha (the
  // alpha of Projection): x_local = K_c^+ (f - Bf^T lambda) - Rf alpha per subdomain
  // (packed local) and u its scatter to the global DOFs; either may be null
  void recover(const double* f_local, const double* lam, const double* alpha, double* x_local, double* u);
  void get_Rf(std::vector<double>& Rf) const;  // packed local, n x 6 column major per subdomain: R_d H_d
  // Projection on the caller's device vectors (x, Px: nlam; alpha: 6 per cluster, the
  // reference's alpha, may be null) with no allocation: the 12 doubles per subdomain
  // of partial sums live in the library's file-scope scratch array (noalloc.hip).
  // stream null: the library's stream, synchronised before return; otherwise the
  // work is only enqueued on `stream`.  Calls that overlap on one GPU race on the
  // shared scratch.
  void project_noalloc(const double* x, double* Px, double* alpha, hipStream_t stream);
  int64_t nlam_c() const;                      // cluster constraints, all clusters

  // One load case: body-force load assembled on the device, or the given packed
  // local loads (sum over subdomains in l2g order).
  void solve(double eps, int max_iter, const double* f_local, double* u_global, double* lam, SolveInfo* info);

  // Time the dominant kernel (gather + GEMV streaming K+_bb (which=0) or S_i
  // (which=1)) with HIP events on the library stream; returns ms per launch and
  // the algorithmic bytes one launch moves.
  double time_stream(int which, int reps, double* bytes);
  // Time the tile GEMM of the block solve (split-FP16 or FP32 format; the kernel
  // alone, after an untimed call that sets the column scales) over 64 columns (which = 0: the
  // deflated K+_bb tiles, 1: S_i) with HIP events on the library stream; returns ms
  // per launch and the algorithmic flops of one launch (2 * 64 * sum n_s^2).
  // which = 2: a plain device copy of known size (*flops = bytes read = written).
  double time_block_gemm(int which, int reps, double* flops);
  // mean device time of the mixed-precision K+_bb GEMM (the band kernel and its
  // reduction, the row form or the FP32 tile GEMM, as configured) over its launches
  // inside one eager mixed-precision block solve of k seeded face tractions
  double time_gemm_in_solve(int k, int* nlaunch);

  // Block of k load cases (column major: column j of f_local at j*packed_n(), of u
  // at j*3*nPoints, of lam at j*nlam).  precision 64: F_i and S_i applied in FP64;
  // 32: F_i and S_i applied in FP32 on the matrix cores with FP32 accumulation,
  // everything else FP64, and each column's stopping test certified by its FP64
  // true residual.  Per column: iterations, status (2 converged, 3 limit, 6 stopped
  // on a non-finite residual; the other columns are unaffected) and the final FP64
  // residual sqrt(g'PMPg)/sqrt(g0'PMPg0).
  void solve_block(double eps, int max_iter, int k, int precision, const double* f_local, double* u, double* lam,
                   int* iters, int* status, double* res, BlockInfo* info);
  // nb blocks of k load cases solved concurrently, one block context each (f_local:
  // nb*k packed local loads; u: nb*k displacements); info->solve_ms is the device
  // time span of all of them, the other times those of the first block
  void solve_blocks(int nb, double eps, int max_iter, int k, int precision, const double* f_local, double* u,
                    int* iters, int* status, double* res, BlockInfo* info);
  int blk_contexts(int n);  // make n block contexts available (context 0 is the engine's own)
  void use_ctx(int c);      // bring block context c into the engine
  // op 0 F, 1 M (preconditioner), 2 P (projector) applied to k dual vectors
  // (column major, ld nlam); precision selects the F_i / S_i arithmetic
  void apply_block(int op, int k, int precision, const double* x, double* y);
  size_t block_bytes() const;

  // Slot lifecycle of a block (continuous batching): blk_begin sets the block
  // width k and empties every slot; blk_join starts load cases in free slots
  // (f_local: one packed local load per slot, or null for the body force);
  // blk_sweeps enqueues PCPG iterations; blk_sync downloads the per-slot scalars;
  // blk_finish recovers the displacements of stopped slots (one 3*nPoints vector
  // per slot into u) and frees them.
  void blk_begin(int k, double eps, int max_iter, int precision);
  void blk_join(const std::vector<int>& slots, const double* f_local);
  void blk_sweeps(int n);
  void blk_sync(BScalars* out);
  void blk_finish(const std::vector<int>& slots, double* u, double* lam, const BScalars* state = nullptr);
  // the same in parts: loads to HBM / start, and recovery / results to the host
  void blk_load(const std::vector<int>& slots, const double* f_local);
  void blk_init(const std::vector<int>& slots);
  void blk_recover(const BScalars* state = nullptr);  // state: the scalars of the last blk_sync
  void blk_collect(const std::vector<int>& slots, double* u, double* lam);
  int blk_width() const { return blk_k; }
  // free slots without recovering them (a cancelled case); they take new cases later
  void blk_release(const std::vector<int>& slots) { blk_set_status(slots, 4); }
  // seeded face-traction loads (traction_loads, problem.cpp) of load cases `cases`
  // into the given slots: the smoothed face fields on the host, the lumping on the
  // device (no packed local vector crosses PCIe)
  void blk_load_traction(const std::vector<int>& slots, uint64_t seed, const std::vector<int>& cases);
  // the recovered displacements of slot j (3 * nPoints, after blk_recover) to device
  // memory dst, on the block's stream
  void blk_copy_u(int j, double* dst);
  // autotuner result of the block GEMM (FP32 format: output row tiles per
  // workgroup 1 or 2; split-FP16 format: K chunk depth 32 or 64 as 1 or 2) and the
  // measured ms of each candidate
  int tuned_rows() const { return tune_rt32; }
  double tuned_ms(int rt) const { return tune_ms32[rt - 1]; }  // rt 3: the band kernel
  bool band_gemm() const { return gemm16 && gemm_band; }
  // runtime settings: hip_graph (1/0: capture and replay the iteration chunks),
  // refine_tol (mixed-precision refinement factor), poll_chunk (iterations
  // between host polls), gemm_rows (1, 2 or auto: block GEMM tile shape, see tuned_rows)
  void configure(const std::string& key, const std::string& value);
  hipStream_t stream() const { return st; }

  // Accessors for tests: per-subdomain results of the setup (host copies).
  void get_R(std::vector<double>& R) const;       // packed local, n x 6 column major per subdomain
  void get_fix(std::vector<int>& fix) const;      // 6 local DOFs per subdomain
  void get_defect(std::vector<int>& d) const;
  void get_load(std::vector<double>& f) const;    // packed local body-force loads
  int64_t packed_n() const { return off_n.back(); }
  std::vector<int64_t> off_n, off_b;

 private:
  std::unique_ptr<Problem> P;
  int device = 0;
  hipStream_t st = nullptr;
  SetupInfo sinfo;
  int maxn = 0, maxns = 0, maxnr = 0, maxnlc = 0;
  int64_t NB = 0;
  // per-subdomain meta (host copies)
  std::vector<int> h_n, h_nr, h_ns, h_nbs, h_nss, h_ldb, h_cl, h_nlc;
  std::vector<int64_t> off_K, off_L, off_R, off_Rb, off_C, off_lc, off_P, off_part, off_r;
  int64_t NR = 0;       // interior rows, all subdomains
  int maxTL = 0, maxLtiles = 0;  // tiles per side / per subdomain of L_rr
  // device meta
  DVec<int> d_n, d_nr, d_ns, d_nbs, d_ldb, d_cl, d_nlc, d_ldn, d_ones, d_sixes, d_zeros;
  DVec<int64_t> d_off_P, d_off_part;
  DVec<int64_t> d_off_n, d_off_b, d_off_K, d_off_L, d_off_R, d_off_Rb, d_off_C, d_off_lc, d_alpha_idx, d_xboff,
      d_v6off;
  // interior factor L_rr (packed lower tiles, inverted diagonal tiles) and the sparse
  // coupling K_br (rows: b positions) / K_rb (rows: interior positions)
  DVec<double> d_Lt, d_brv, d_rbv;
  DVec<int64_t> d_brp, d_rbp;
  DVec<int> d_brc, d_rbc, d_bloc, d_rsub, d_rloc;
  DVec<int> d_perm1, d_gdof, d_gptr, d_gidx, d_lcmap, d_bsub;
  DVec<int> d_gpos;  // b position (within its subdomain) of each gidx entry
  DVec<double> d_Kt, d_St, partbuf, xb;  // packed lower tiles of K+_bb and S_i, tile partials
  int maxT = 0, maxtiles = 0;
  DVec<unsigned char> d_owner;
  DVec<double> d_kdiag, d_f, d_K, d_S, d_R, d_Rb, d_C, d_BcT, d_H, d_invGtG;
  DVec<int> d_fixpos, d_defect, d_err;
  // dual rows
  DVec<int> d_ra, d_rb, d_rsa, d_rsb, d_rpa, d_rpb;
  DVec<double> d_wgt;
  // clusters
  std::vector<int> h_nlcl;
  std::vector<int64_t> h_la_off, h_Aoff;
  std::vector<int> h_lcmap, h_gcsrc;  // host copies of d_lcmap, d_gc_src (saved state)
  DVec<int> d_gc_src, d_nlcl, d_nrcl, d_d0, d_nlen;
  DVec<int64_t> d_gc_off, d_la_off, d_Aoff;
  DVec<double> d_Ainv;
  int max_cl_len = 0;
  // work vectors
  DVec<double> lam, g, w, Fw, Pg, z, Pz, dvec, yb, cvec, avec, gc, la, v6, h, hg, alpha, part, X1, X2, rtmp, vloc, xloc,
      ug;
  DVec<Scalars> sc;
  int nparts = 0;
  bool lumped = false;
  std::vector<int> h_fix, h_def;

  // ---- block state (allocated on the first block call)
  bool use_graph = true;
  // in-solve GEMM probe (time_gemm_in_solve): events around every K+_bb GEMM launch
  Events* gemm_probe = nullptr;
  size_t gemm_probe_n = 0;
  int tune_rt32 = 1, gemm_rows_forced = 0;
  int gemm_lds_pad = 0, gemm_exact = 0, gemm_pf = 1;  // split-FP16 GEMM: residency cap, tile order, prefetch
  // split-FP16 GEMM reading every tile once (bands of row blocks, partial sums in
  // b_p16 from d_p16off[d]) instead of once per row block
  bool gemm_band = true;
  double tune_ms32[3] = {0, 0, 0};  // candidates 1, 2 (tuned_rows) and the band kernel
  int poll_chunk = 8;
  hipGraphExec_t gx_plain = nullptr, gx_blk = nullptr;
  int gx_plain_n = -1;
  int64_t gx_blk_key = -1;
  void run_graph(hipGraphExec_t* exec, const std::function<void()>& enqueue);
  bool blk_ready = false;
  int blk_k = 0, blk_prec = 64;
  void blk_set_status(const std::vector<int>& slots, int status);
  void blk_enqueue_sweeps(int n);
  void autotune();
  // b_yb32 = A b_xb32 on the matrix cores, A = K+_bb (which 0) or S_i (which 1) in
  // the block format (split FP16 or FP32)
  void gemm_lo(int which, const int* gate);
  int maxTb = 0, maxTr = 0, maxTcr = 0, maxTcl = 0;
  int sAc = 1;  // split-K parts of the Ac^-1 block product
  int64_t NBP = 0, NRP = 0, CRP = 0, GCP = 0;
  DVec<int64_t> b_obrow, b_orrow, b_xoff, b_roff, b_croff, b_cro, b_gcrow, b_gcoff, b_lcrow, b_lcpos, b_arow, b_abase_cr,
      b_v6off, b_v6row, b_crrow;
  DVec<int> b_crn, b_qrow;
  DVec<double> b_CR, b_CRt, b_Gr, b_kpart, b_gt;
  DVec<int> b_ghsub, b_ghslot;  // kb_gh: subdomains staged per 64 dual rows, slots per row
  DVec<int64_t> b_crto;
  DVec<float> b_Kt32, b_St32, b_xb32, b_yb32;
  // split-FP16 block GEMM (option block_gemm=f16x2): b_Kt32/b_St32 hold hi and lo FP16
  // planes of A_d / s_d, b_sK/b_sS the power-of-two s_d, b_sx the per-column scales of X
  bool gemm16 = true;
  DVec<float> b_sK, b_sS, b_sx, b_p16;
  DVec<int64_t> d_p16off;
  int64_t p16_n = 0;
  DVec<int> b_xob, b_xor;  // XCD orders of the tile GEMMs over b rows (n_s^2) and interior rows (n_r^2)
  int n_xob = 0, n_xor = 0;
  DVec<int> b_gcsub, b_gcp0, b_gcbeg;  // chunks of b positions for G^T x (kb_gtx)
  DVec<int> b_srow;  // b-block rows of the + and - entry of every dual row (kb_scatter)
  DVec<int> d_face_meta;                // subdomains on the x = 1 face (traction lumping)
  DVec<int64_t> d_face_off;
  DVec<double> d_trac;                  // face traction fields of the loads being joined
  int n_face = -1;
  int n_gchunk = 0;
  DVec<double> pgn;    // |P g0|^2 and |P g|^2 of the plain solve
  DVec<double> b_gtp;  // G^T x chunk partials
  DVec<double> b_xb, b_yb, b_cr, b_coef, b_v6, b_gc, b_la, b_h, b_hg, b_hgf, b_lam, b_g, b_w, b_Fw, b_Pg, b_z, b_Pz, b_d, b_g64, b_part,
      b_Vr, b_Xr, b_Tr, b_V, b_V2, b_X, b_U;
  DVec<BScalars> b_sc;
  int b_nparts = 0;
  void block_prepare();
  void bF(int k, int prec, const double* win, double* out, const double* sub, const double* dotw,
          const int* gate = nullptr);
  void bP(int k, const double* x, double* out, double* hbuf, const double* dotv, const int* gate = nullptr);
  void bM(int k, int prec, const double* x, double* out, const int* gate = nullptr);
  void bPM(int k, int prec, const double* x, double* out, double* hbuf, const int* gate = nullptr);
  void b_gtx_coarse(int k, const double* x, double* hbuf, const int* gate);
  void b_corr(int k, int with_R, int defl, const int* gate, const float* yf = nullptr);
  void crq_products(const int* gate);
  void b_kplus(int k, bool interior_too, double* V);  // K_c^+ pieces of the local block V
  void b_rr_solve(double* X);
  std::vector<std::unique_ptr<BlockCtx>> ctxs_;  // ctxs_[i]: context i while it is not the current one
  int cur_ctx_ = 0;
  void swap_ctx(BlockCtx& c);
  void alloc_ctx_buffers();

  void build_meta();
  void factor_subdomains();
  void build_bc_and_coarse();
  void coarse_layout();  // constraint offsets and device copies from h_nlc, h_nlcl
  void alloc_work();     // solve work vectors (end of set-up and of a state load)
  void state_fields(StateStream& s, hipStream_t cs);
  void partial_cholesky(double* M, const std::vector<int64_t>& moff, const std::vector<int>& ld,
                        const std::vector<int>& start, const std::vector<int>& ncols, const std::vector<int>& end);
  // solve pieces (device vectors)
  void F_apply(const double* w_in, double* out, const double* sub_from, const double* dotw, const int* done);
  void proj(const double* x, double* out, double* hbuf, double* alpha_out, const double* dotv, const int* done);
  void M_apply(const double* x, double* out, const int* done);
  void symv(const double* tiles, const double* x, const int64_t* xoff, const int* done);  // yb = A x
  void prep_t(const double* vloc_in);        // X1 = [v_r; t_b], t_b = v_b - K_br K_rr^-1 v_r
  void kplus_b_from_t();                     // yb = K+_bb t_b = (K^+ v)_b
  void cluster_solve_t(const double* vloc_in);  // la = Ac^-1 [C^T t; -R^T v]
  void interior(double* xout);               // xout = [K_rr^-1 v_r - K_rr^-1 K_rb yb ; yb] in local order
  // X <- K_rr^-1 X = W^T W X on the interior rows of every subdomain (element (i, j)
  // of d at X[xbase[d] + i*rs + j*cs]); xbase holds the subdomains' bases from sub0 on
  void recover_u(const double* f, const int* gdof, const unsigned char* owner, double* u);
  void coarse_to_sub(const double* a, double sgn, std::vector<double>& hv) const;
  void sub_to_coarse(const std::vector<double>& v, double* e) const;
  void rr_solve(int nsub, int sub0, int k, double* X, const int64_t* xbase, int64_t rs, int64_t cs, const int* csv,
                double* tmp, const int64_t* tbase);  // tmp: same layout as X
};

}  // namespace hfls
