// Launch wrappers of the gfx950 kernels (kernels.hip, kernels2.hip).  Every
// wrapper enqueues on the given stream and returns the launch status.  Arrays
// named *off hold per-matrix (per-subdomain) offsets into packed buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>
#include <cstdint>

namespace hfls {

struct PanelDesc {
  int k0, w, end, active;  // panel columns [k0,k0+w), trailing rows/cols up to end
};

struct SyrkDesc {  // update by columns [k0,k0+w) of rows/cols t0 <= c <= r < end, c < cend
  int k0, w, t0, cend, end, active;
};

struct AsmSub {  // box subdomain for element assembly
  int e0[3], ne[3], n, nr;  // nr > 0: also write the identity coupling M[n+i][i] = 1, i < nr
  double E;
  // per-element material (mn[0] > 0): an element in geometric partition
  // pid = gi/mn[0] + (gj/mn[1]) mN[0] + (gk/mn[2]) mN[0] mN[1] (global element indices)
  // has E when pid is even and E1 when odd (Mesh.cpp:271-288, Data.cpp:47-48)
  int mn[3], mN[2];
  double E1;
};

struct Scalars {  // PCPG scalars kept on the device
  double rho, gamma, gPz, gPz_prev, wFw, norm0, eps, gPg0;
  int it, done, max_iter, err;
};

// ---------------------------------------------------------------- interior solves (lrr.hip)
hipError_t launch_lrr_diag_inv(hipStream_t st, int nsub, int maxT, const int* nr, double* Lt, const int64_t* loff);
hipError_t launch_lrr_trtri(hipStream_t st, int nsub, int maxT, const int* nr, const double* Lt, const int64_t* loff,
                            double* Wt, const int64_t* woff);
hipError_t launch_trmv(hipStream_t st, int nsub, int maxT, int op, int k, const int* nr, const double* Wt,
                       const int64_t* woff, const double* X, const int64_t* xbase, int64_t rs, int64_t cs,
                       double* Y, const int64_t* ybase, const int* csv = nullptr);
hipError_t launch_csr_apply(hipStream_t st, int64_t row0, int64_t row1, int sub0, int k, const int64_t* ptr,
                            const int* col, const double* val, const int* rsub, const int* rloc, const double* X,
                            const int64_t* xbase, int64_t xrs, int64_t xcs, double* Y, const int64_t* ybase,
                            int64_t yrs, int64_t ycs, const int* csv, double alpha, double beta);
hipError_t launch_csr_extract(hipStream_t st, int64_t row0, int64_t row1, int sub0, int trans, const int64_t* ptr,
                              const int* col, const int* rsub, const int* rloc, const int* nr, const double* M,
                              const int64_t* moff, const int* ld, double* val);
hipError_t launch_rr_axpby(hipStream_t st, int nsub, int maxnr, int k, const int* nr, const double* A,
                           const int64_t* abase, double* B, const int64_t* bbase, int64_t rs, int64_t cs,
                           double alpha, double beta);

// ---------------------------------------------------------------- setup
hipError_t launch_assemble(hipStream_t st, int nmat, const AsmSub* subs, const int64_t* moff, const int* ld,
                           const int* pinv, const int64_t* pinv_off, double* M, double* f, const int64_t* f_off,
                           double* kdiag, double dx, double dy, double dz, double nu, double ax, double ay,
                           double az);
hipError_t launch_panel(hipStream_t st, int nmat, int maxrows, const PanelDesc* desc, const int64_t* moff,
                        const int* ld, double* M, double* Winv, int* err);
hipError_t launch_syrk(hipStream_t st, int nmat, int64_t maxtiles, const SyrkDesc* desc, const int64_t* moff,
                       const int* ld, double* M);
hipError_t launch_copy_block(hipStream_t st, int nmat, int maxn, const double* src, const int64_t* soff,
                             const int* sld, const int* srow0, const int* n, double* dst, const int64_t* doff,
                             const int* dld, double scale, int aug);
hipError_t launch_detect(hipStream_t st, int nmat, const double* M, const int64_t* moff, const int* ld,
                         const int* row0, const int* m, const double* kdiag, const int64_t* kd_off,
                         const int* perm, const int* prow0, double* Rs, const int64_t* rs_off, const int* ldr,
                         int* defect);
hipError_t launch_coldot(hipStream_t st, int nmat, int maxcols, const double* L, const int64_t* loff,
                         const int* ld, const int* row0, const int* mrows, const int* ncols, const double* Y,
                         const int64_t* yoff, const int* ldy, double* out, const int64_t* ooff, const int* ldo,
                         int nrhs, double alpha, double beta);
hipError_t launch_bwd(hipStream_t st, int nmat, const double* L, const int64_t* loff, const int* ld,
                      const int* m, double* X, const int64_t* xoff, const int* ldx, int nrhs);
hipError_t launch_fwd(hipStream_t st, int nmat, const double* L, const int64_t* loff, const int* ld,
                      const int* m, const int* mall, double* X, const int64_t* xoff, const int* ldx, int nrhs);
hipError_t launch_gs_pivots(hipStream_t st, int nmat, double* R, const int64_t* roff, const int* n,
                            const int* perm, const int64_t* perm_off, const int* nr, const int* ns,
                            const int* nbs, double* Rb, const int64_t* rb_off, double* Rblk,
                            const int64_t* rblk_off, const double* kdiag, const int64_t* kd_off, double* M2,
                            const int64_t* m2off, const int* ld2, int* fixpos);
hipError_t launch_permute(hipStream_t st, int nsub, int maxlen, int scatter, const int* len, const int* perm,
                          const int64_t* poff, const double* src, const int64_t* soff, double* dst,
                          const int64_t* doff, int nrhs, int64_t sstride, int64_t dstride);
hipError_t launch_copy_cols(hipStream_t st, int nmat, int maxcols, const int* rows, const int* cols,
                            const double* src, const int64_t* soff, const int* lds, double* dst,
                            const int64_t* doff, const int* ldd, double scale);
hipError_t launch_scale_lower(hipStream_t st, int nmat, int maxcols, const int* row0, const int* n, double* M,
                              const int64_t* moff, const int* ld, double s);
hipError_t launch_gemm_nn(hipStream_t st, int nmat, int maxn, int maxcol, const int* n, const int* nin,
                          const int* ncol, const double* A, const int64_t* aoff, const int* lda, const double* X,
                          const int64_t* xoff, const int* ldx, double* Y, const int64_t* yoff, const int* ldy,
                          double alpha);
// y_i = A_i x1_i + beta * B_i x2_i  (A_i: m x m1, B_i: m x m2, column major; x staged in LDS)
hipError_t launch_gemv_two(hipStream_t st, int nsub, int maxm, int maxlds, const int* m, const int* m1,
                           const int* m2, const double* A, const int64_t* aoff, const int* lda, const double* B,
                           const int64_t* boff, const int* ldb, const double* x1, const int64_t* x1off,
                           const double* x2, const int64_t* x2off, double beta, double* y, const int64_t* yoff);
hipError_t launch_tdots(hipStream_t st, int nmat, int maxpairs, const int* np, const int* ni, const int* nj,
                        const double* X, const int64_t* xoff, const int* ldx, const double* Y, const int64_t* yoff,
                        const int* ldy, double* out, const int64_t* ooff, const int* ldo, double alpha);

// ---------------------------------------------------------------- solve
hipError_t launch_gemv_gather(hipStream_t st, int nsub, int maxrows, const double* A, const int64_t* aoff,
                              const int* lda, const int* ns, const int64_t* boff, const int* gptr,
                              const int* gidx, const double* v, const double* wgt, const double* xb,
                              const int64_t* xboff, double* y, const int* done);
hipError_t launch_corr_dots(hipStream_t st, int nsub, int maxns, const double* C, const int64_t* coff,
                            const int* ldc, const int* nlc, const int64_t* lcoff, const double* Rb,
                            const int64_t* rboff, const int* ns, const int64_t* boff, const int* gptr,
                            const int* gidx, const double* v, const double* xb, double* cvec, double* avec,
                            const int* done);
hipError_t launch_cluster_gc(hipStream_t st, int ncl, int maxlen, const int* gc_src, const int64_t* gc_off,
                             const int* nlcl, const int* nrcl, const int* d0, const double* cvec,
                             const double* avec, double* gc, const int64_t* la_off, const int* done);
hipError_t launch_cluster_gemv(hipStream_t st, int ncl, int maxlen, const double* Ainv, const int64_t* aoff,
                               const int* nlen, const double* gc, const int64_t* la_off, double* la,
                               const int* done);
hipError_t launch_corr_apply(hipStream_t st, int nsub, int maxrows, const double* C, const int64_t* coff,
                             const int* ldc, const int* nlc, const int64_t* lcoff, const int* lcmap,
                             const double* Rb, const int64_t* rboff, const int* ns, const int64_t* boff,
                             const int64_t* alpha_idx, const double* la, double* y, const int* done);
// column dots with one wave per output: out[o + i] = alpha * sum_j A[j + i*n] x[xo + j]  (A symmetric, n x n)
hipError_t launch_sym_gemv_waves(hipStream_t st, int ncl, int maxlen, const double* A, const int64_t* aoff,
                                 const int* nlen, const double* x, const int64_t* xoff, double* out,
                                 const int* done);
hipError_t launch_scatter(hipStream_t st, int64_t nlam, const int* ra, const int* rb, const double* y,
                          const double* wgt, const double* sub_from, double* out, const double* dotw,
                          double* part, const int* done);
hipError_t launch_reduce(hipStream_t st, int nparts, const double* part, Scalars* sc, int mode);
hipError_t launch_axpy2(hipStream_t st, int64_t nlam, double* lam, double* g, const double* w,
                        const double* Fw, const Scalars* sc);
hipError_t launch_proj_dots(hipStream_t st, int nsub, const double* Rb, const int64_t* rboff, const int* ns,
                            const int64_t* boff, const int* gptr, const int* gidx, const double* x,
                            double* v, const int* done);
hipError_t launch_proj_coarse(hipStream_t st, int nsub, int ncol, const double* v, const double* H,
                              const int* cl, const double* invGtG, double* h, double* alpha, double sgn,
                              const int* done, double asgn = 1.0);
hipError_t launch_gh(hipStream_t st, int64_t nlam, const int* rsub_a, const int* rsub_b, const int* rpos_a,
                     const int* rpos_b, const double* Rb, const int64_t* rboff, const int* ns, const double* h,
                     const double* x, double* out, double sgn, const double* dotv, double* part,
                     const int* done);
hipError_t launch_update_w(hipStream_t st, int64_t nlam, double* w, const double* Pz, const Scalars* sc);
hipError_t launch_sub_bt(hipStream_t st, int nsub, int maxns, const int* nr, const int* ns, const int* perm,
                         const int64_t* noff, const int64_t* boff, const int* gptr, const int* gidx,
                         const double* lam, const double* BcT, const int64_t* coff, const int* ldc, const int* nlc,
                         const int64_t* lcoff, const int* lcmap, const double* la, double* v);
// f (one packed local load, zeroed before) += the lumped face traction t of one load
// case (traction_loads' lumping) over the nface subdomains on the x = 1 face
hipError_t launch_traction_lump(hipStream_t st, int nface, const int* meta, const int64_t* foff, int ny1, int nz1,
                                double hy, double hz, const double* t, double* f);
// device address of the file-scope scratch array of the allocation-free entry
// point on the current device (looked up once per device), or an error if
// `doubles` exceeds its size (NOALLOC_SCRATCH_DOUBLES)
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = 196608;  // 1.5 MiB: 12 doubles for each of 16384 subdomains
double* noalloc_scratch(int device, int64_t doubles);

hipError_t launch_recover(hipStream_t st, int nsub, int maxn, const int* n, const int64_t* noff, const double* R,
                          const int64_t* roff, const int64_t* alpha_idx, const double* la, const double* h,
                          const double* x, const int* gdof, const unsigned char* owner, double* u);

// ---------------------------------------------------------------- symmetric tiles (symv.hip)
hipError_t launch_pack_tiles(hipStream_t st, int nsub, int maxtiles, const int* ns, const double* A,
                             const int64_t* aoff, const int* lda, double* P, const int64_t* poff);
hipError_t launch_gather(hipStream_t st, int64_t nb, const int* gptr, const int* gidx, const double* v,
                         const double* wgt, double* xb, const int* done);
hipError_t launch_symv_tiles(hipStream_t st, int nsub, int maxT, const int* ns, const double* P, const int64_t* poff,
                             const double* x, const int64_t* xoff, double* part, const int64_t* partoff,
                             const int* done);
hipError_t launch_symv_reduce(hipStream_t st, int64_t nb, const int* bsub, const int* ns, const int64_t* boff,
                              const double* part, const int64_t* partoff, double* y, const int* done);

// out[0] = |x|^2 (one workgroup, deterministic order)
hipError_t launch_sqnorm(hipStream_t st, int64_t n, const double* x, double* out);

// ---------------------------------------------------------------- tile GEMM (tgemm.hip)
// The tile GEMMs put all row blocks of a subdomain on one XCD: entry 8 g + x of an
// XCD order (length norder, a multiple of 8) is the g-th subdomain of XCD x, -1 for
// none.  Without one, subdomain d goes to XCD d % 8.  xcd_order() balances the
// XCDs' work (longest first, each to the least loaded XCD).
std::vector<int> xcd_order(const std::vector<double>& work);
// Y_d = alpha op(A_d) X_d (+ Z_d); A_d: mode 0 symmetric lower 64x64 tiles, mode 1
// general tiles, mode 2 compact column major (lda); X, Y, Z row major, ld kp.
hipError_t launch_tgemm_f32(hipStream_t st, int nsub, int maxRo, int mode, int op, const int* mrows, const int* mcols,
                            const float* A, const int64_t* aoff, const int* lda, const float* X, const int64_t* xoff,
                            double* Y, const int64_t* yoff, const double* Z, int kp, float alpha, const int* gate,
                            int rt = 1,              // rt: 64-row output tiles per workgroup (1 or 2)
                            float* Yf = nullptr,     // Yf: store the FP32 result there instead of Y (Z unused)
                            const int* xcd_order = nullptr, int norder = 0);  // see xcd_order()
hipError_t launch_tgemm_f64(hipStream_t st, int nsub, int maxRo, int mode, int op, const int* mrows, const int* mcols,
                            const double* A, const int64_t* aoff, const int* lda, const double* X,
                            const int64_t* xoff, double* Y, const int64_t* yoff, const double* Z, int kp,
                            double alpha, const int* gate, int S = 1, double* Ypart = nullptr,  // S: split K
                            const float* Zf = nullptr,  // Zf: an FP32 addend in place of Z
                            const int* xcd_order = nullptr, int norder = 0);  // see xcd_order()
// Y = A C (+ Z or + Zf) for tiled A of at most 64 columns (one tile column), C row major
hipError_t launch_corr_f64(hipStream_t st, int nsub, int maxRo, const int* mrows, const int* mcols, const double* A,
                           const int64_t* aoff, const double* C, const int64_t* coff, double* Y, const int64_t* yoff,
                           const double* Z, const float* Zf, int kp, const int* gate, const int* order, int norder);
hipError_t launch_pack_f32(hipStream_t st, int nsub, int maxtiles, int mode, const int* mrows, const int* mcols,
                           const double* A, const int64_t* aoff, const int* lda, float* P, const int64_t* poff);
hipError_t launch_pack_f64(hipStream_t st, int nsub, int maxtiles, int mode, const int* mrows, const int* mcols,
                           const double* A, const int64_t* aoff, const int* lda, double* P, const int64_t* poff);
// matrix-core ceiling on register data with random operands (no memory traffic): nwg workgroups of 4
// waves, each wave `iters` rounds of independent MFMAs; one value per thread to out
// (nwg * 256 doubles).  f64: v_mfma_f64_16x16x4f64, else v_mfma_f32_32x32x2f32.
// Returns the flops of the launch in *flops.
hipError_t launch_mfma_ceiling(hipStream_t st, bool f64, int nwg, int iters, double* out, double* flops);

// ---------------------------------------------------------------- block PCPG (block.hip)
constexpr int KMAX = 64;  // load cases per block
struct BScalars {         // per-column PCPG scalars of a block solve
  double rho[KMAX], gamma[KMAX], gPz[KMAX], gPz_prev[KMAX], wFw[KMAX], norm0[KMAX], res[KMAX], ref[KMAX];
  int it[KMAX], done[KMAX], repl[KMAX], ncert[KMAX];
  int hgok[KMAX];          // the saved H alpha of the column's certified true residual is valid
  int certnow[KMAX];       // certified by the last certification reduction
  int waitn[KMAX];         // gamma reductions the column has spent waiting for its FP64 evaluation
  int certified;           // some column was certified since the last gamma reduction
  double eps, rtol;        // stopping tolerance; refinement factor (0: none)
  int max_iter, err;
  int cert;                // 1: stopping tests are confirmed on the FP64 true residual
  int pending;             // some column waits for an FP64 true-residual evaluation
  int active;              // some column is still iterating
  int restart;             // some column restarts from its FP64 true residual
  double colsum[KMAX];     // column sums of the last reduction
};

hipError_t launch_kb_gather(hipStream_t st, int64_t nb, int k, int kp, const int* qrow, const int* gptr, const int* gidx, const double* v,
                            const double* wgt, double* xb, float* xb32, const int* gate = nullptr);
// G^T x per subdomain (out rows 6d + m) through chunk partials (part, 6 nchunk rows);
// chunk c covers positions cp0[c] .. cp0[c] + 255 of subdomain csub[c], subdomain d
// owns chunks cbeg[d] .. cbeg[d+1] - 1
hipError_t launch_kb_gtx(hipStream_t st, int nsub, int nchunk, int k, int kp, const int* ns, const int* csub,
                         const int* cp0, const int* cbeg, const int64_t* boff, const int* gptr, const int* gidx,
                         const int* gpos, const double* Rb, const int64_t* offRb, const double* x, double* part,
                         double* out, const int* gate = nullptr);
hipError_t launch_kb_to_f32(hipStream_t st, int64_t n, const double* a, float* b);
hipError_t launch_kb_copy(hipStream_t st, int64_t n, const double* a, double* b);
hipError_t launch_kb_rtf(hipStream_t st, int nsub, int k, int kp, int64_t npacked, const int* n, const int64_t* noff,
                         const double* R, const int64_t* roff, const double* V, double* a);
hipError_t launch_kb_cluster_gc(hipStream_t st, int ncl, int maxlen, int k, int kp, const int* gc_src,
                                const int64_t* gc_off, const int* nlcl, const int* nrcl, const int* d0,
                                const double* cr, const int64_t* lcpos, const double* a, const int64_t* abase,
                                double* gc, const int64_t* gcrow, const int* gate = nullptr);
hipError_t launch_kb_coef(hipStream_t st, int nsub, int maxrows, int k, int kp, const int* nlc, const int64_t* lcoff,
                          const int64_t* lcrow, const int64_t* arow, const double* la, int with_R, int defl,
                          const double* cr, const int64_t* crow, double* coef, const int* gate = nullptr);
hipError_t launch_kb_cols(hipStream_t st, int nsub, int maxns, int dir, int ncol, int kp, const int* ns,
                          const int64_t* obrow, double* A, const int64_t* aoff, const int* lda, double* X);
// split-FP16 form of packed FP32 tiles, in place (tgemm16.hip): scale[d] = power of
// two mapping max |A_d| into [2^14, 2^15); each tile becomes hi and lo FP16 planes
hipError_t launch_split16(hipStream_t st, int nsub, int maxtiles, const int* ns, const int64_t* poff, float* P32,
                          float* scale);
// Yf_d = alpha A_d X_d, A_d split-FP16 symmetric tiles (mode 0) with scale sA[d], X FP32
// row major (ld kp); sx (nsub * kp) receives the per-column scales of X
hipError_t launch_tgemm_h16(hipStream_t st, int nsub, int maxT, const int* ns, const float* A, const int64_t* aoff,
                            const float* sA, const float* X, const int64_t* xoff, float* sx, int kp, float alpha,
                            const int* gate, float* Yf, const int* xcd_order = nullptr, int norder = 0,
                            int kc = 32,     // kc: K chunk depth staged per step (32 or 64)
                            int scales = 1,
                            int lds_pad = 0,  // unused dynamic LDS per workgroup (caps residency per CU)
                            int exact = 0,    // tile order rotated by the row block itself (else its pair)
                            int pf = 1);      // K chunks in flight while one is multiplied (1 or 2) // scales 0: sx already holds the scales of this X
// The same product reading every stored tile once (tgemm16b.hip): workgroups own
// bands of HFLS_H16B_M row blocks; the transposed contributions to the row blocks
// above a band go through `part` (tgemm_h16b_part_floats(n_s, kp) floats per
// subdomain from poff[d]) and are added in a fixed order.
#define HFLS_H16B_M 4
int64_t tgemm_h16b_part_floats(int ns, int kp);
hipError_t launch_tgemm_h16b(hipStream_t st, int nsub, int maxT, const int* ns, const float* A, const int64_t* aoff,
                             const float* sA, const float* X, const int64_t* xoff, float* sx, int kp, float alpha,
                             const int* gate, float* Yf, const int* xcd_order, int norder, int scales, float* part,
                             const int64_t* poff);
// sx[d * kp + c]: power-of-two scale of column c of X_d (tgemm16.hip)
hipError_t launch_kb_colscale(hipStream_t st, int nsub, const int* ns, const int64_t* xoff, int kp, const float* X,
                              float* sx, const int* gate);
hipError_t launch_kb_deflate_f32(hipStream_t st, int nsub, int maxtiles, const int* ns, const double* P,
                                 const int64_t* poff, const double* CRQ, const int64_t* coff, const int* ldc,
                                 const int* nlc, float* P32);
// srow[2r], srow[2r+1]: b-block rows of dual row r's + and - entries (-1: none)
hipError_t launch_kb_scatter(hipStream_t st, int64_t nlam, int k, int kp, const int* srow, const double* y,
                             const double* wgt, const double* sub, double* out, const double* dotw, double* part,
                             const int* gate = nullptr,
                             const float* yf = nullptr);  // yf: FP32 b-block in place of y
// subdomains whose h rows kb_gh stages in LDS per group of 64 dual rows
constexpr int GH_SLOTS = 8;
hipError_t launch_kb_gh(hipStream_t st, int64_t nlam, int k, const int* rsa, const int* rsb, const int* gsub,
                        const int* rslot, const double* Gr, const double* h, const double* x, double* out, double sgn,
                        const double* dotv, double* part, const int* gate = nullptr);
hipError_t launch_kb_proj_coarse(hipStream_t st, int nsub, int ncol, int k, int kp, const double* v,
                                 const int64_t* vrow, const double* H, const int* cl, const double* invGtG, double* h,
                                 double* alpha, double sgn, const int* gate = nullptr);
hipError_t launch_kb_reduce(hipStream_t st, int nparts, int k, const double* part, BScalars* sc, int mode,
                            const int* gate = nullptr);
// hgf[q*k + j] = hg[q*k + j] (q < n6) for the columns certified by the last
// certification reduction (skipped on the device when none was)
hipError_t launch_kb_save_hg(hipStream_t st, int n6, int k, BScalars* sc, const double* hg, double* hgf);
hipError_t launch_kb_update(hipStream_t st, int64_t nlam, int k, int mode, double* a, double* b, const double* c,
                            const double* e, const BScalars* sc, const int* gate = nullptr);
hipError_t launch_kb_permute(hipStream_t st, int nsub, int maxn, int dir, int k, int kp, int64_t npacked,
                             const int* n, const int* nr, const int* perm, const int64_t* noff, double* V, double* Vr,
                             const int64_t* orrow, double* Vb, const int64_t* obrow);
hipError_t launch_kb_sub_bt(hipStream_t st, int nsub, int maxns, int k, int64_t npacked, const int* nr, const int* ns,
                            const int* perm, const int64_t* noff, const int64_t* boff, const int* gptr,
                            const int* gidx, const double* lam, double* V);
hipError_t launch_kb_recover(hipStream_t st, int nsub, int maxn, int k, int kp, int64_t npacked, int64_t nglob,
                             const int* n, const int64_t* noff, const double* R, const int64_t* roff,
                             const int64_t* arow, const double* la, const double* h, const double* X, const int* gdof,
                             const unsigned char* owner, double* U);

}  // namespace hfls
