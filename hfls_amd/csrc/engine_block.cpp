// Block PCPG: k load cases of one mesh solved together (Cluster.cpp:1363-1458 run
// per column, recovery as Cluster.cpp:1648-1678).
//
// The dense local operators become GEMMs over a block of right-hand sides on the
// matrix cores (tgemm.hip): K+_bb and S_i (packed symmetric 64x64 tiles) in FP32
// with FP32 accumulation inside the iteration (precision 32) or FP64 (precision
// 64); the correction blocks [C_i | R_b,i], the cluster inverse, the interior
// blocks K_rr^-1 and K_rr^-1 K_rb of the initialisation and the recovery always in
// FP64.  Dual vectors, dot products, the projector and the coarse problem are FP64.
//
// Mixed precision: the recursive residual g is advanced with the FP32 operator.
// When a column meets the stopping test, its FP64 true residual g = F lam - d is
// evaluated and the test repeated on it (certification); a column that fails has
// its recursive residual replaced by the true one and resumes.  The stopping test
// is checked every `chunk` iterations from the column's own start, so a column's
// iterates do not depend on the other columns of the block.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace hfls {

static constexpr int KP = KMAX;  // leading dimension of row-major blocks
static int64_t r64(int64_t x) { return (x + 63) / 64 * 64; }
static int64_t a32(int64_t x) { return (x + 31) / 32 * 32; }

template <class T>
static void zero_alloc(DVec<T>& v, size_t n, hipStream_t st) {
  v.alloc(std::max<size_t>(n, 1));
  HC(hipMemsetAsync(v.p, 0, v.n * sizeof(T), st));
}

void Engine::block_prepare() {
  if (blk_ready) return;
  const int N = P->nsub, NC = P->ncl;
  const int64_t nl = P->nlam, np = off_n[N], ng = 3 * (int64_t)P->nPoints;
  std::vector<int64_t> obrow(N), orrow(N), xoff(N), roff(N), croff(N), cro(N), cro_rb(N), cro_q(N), cro_y(N),
      arow(N), abase(N), v6row(N), v6off(N);
  std::vector<int> crn(N);
  int64_t rb = 0, rr = 0, rc = 0, co = 0;
  maxTb = maxTr = maxTcr = maxTcl = 0;
  for (int d = 0; d < N; d++) {
    obrow[d] = rb;
    rb += r64(h_ns[d]);
    maxTb = std::max(maxTb, (h_ns[d] + 63) / 64);
    orrow[d] = rr;
    rr += r64(h_nr[d]);
    maxTr = std::max(maxTr, (h_nr[d] + 63) / 64);
    crn[d] = h_nlc[d] + 18;
    croff[d] = rc;
    rc += r64(crn[d]);
    maxTcr = std::max(maxTcr, (crn[d] + 63) / 64);
    cro[d] = co;
    cro_rb[d] = co + (int64_t)h_nlc[d] * h_ldb[d];
    cro_q[d] = cro_rb[d] + 6 * (int64_t)h_ldb[d];
    cro_y[d] = cro_q[d] + 6 * (int64_t)h_ldb[d];
    co += a32((int64_t)h_ldb[d] * crn[d]);
    xoff[d] = obrow[d] * KP;
    roff[d] = orrow[d] * KP;
    v6row[d] = 64 * (int64_t)d;
    v6off[d] = v6row[d] * KP;
  }
  NBP = rb;
  {  // XCD orders balancing the tile GEMMs' work over the 8 XCDs
    std::vector<double> wb(N), wr(N);
    for (int d = 0; d < N; d++) {
      wb[d] = (double)h_ns[d] * h_ns[d];
      wr[d] = (double)h_nr[d] * h_nr[d];
    }
    const std::vector<int> ob = xcd_order(wb), orr = xcd_order(wr);
    b_xob.upload(ob, st);
    b_xor.upload(orr, st);
    n_xob = (int)ob.size();
    n_xor = (int)orr.size();
  }
  {  // partial sums of the band form of the split-FP16 GEMM, per subdomain
    std::vector<int64_t> po(N);
    p16_n = 0;
    for (int d = 0; d < N; d++) {
      po[d] = p16_n;
      p16_n += tgemm_h16b_part_floats(h_ns[d], KP);
    }
    d_p16off.upload(po, st);
  }
  {  // chunks of at most 256 b positions for G^T x
    std::vector<int> csub, cp0, cbeg(N + 1, 0);
    for (int d = 0; d < N; d++) {
      for (int p = 0; p < h_ns[d]; p += 256) {
        csub.push_back(d);
        cp0.push_back(p);
      }
      cbeg[d + 1] = (int)csub.size();
    }
    b_gcsub.upload(csub, st);
    b_gcp0.upload(cp0, st);
    b_gcbeg.upload(cbeg, st);
    n_gchunk = (int)csub.size();
  }
  NRP = rr;
  CRP = rc;
  std::vector<int64_t> gcrow(NC), gcoff(NC), croff_el(N);
  int64_t gr = 0;
  for (int cl = 0; cl < NC; cl++) {
    const int nlen = h_nlcl[cl] + 6 * P->nsub_cl;
    gcrow[cl] = gr;
    gcoff[cl] = gr * KP;
    gr += r64(nlen);
    maxTcl = std::max(maxTcl, (nlen + 63) / 64);
  }
  GCP = gr;
  // the Ac^-1 product has few output blocks (maxTcl row blocks per cluster): split
  // its inner dimension so that about 1024 workgroups run, keeping >= 2 K chunks each
  sAc = std::max(1, std::min({8, 1024 / std::max(1, NC * maxTcl), maxTcl}));
  for (int d = 0; d < N; d++) {
    const int cl = h_cl[d];
    croff_el[d] = croff[d] * KP;
    arow[d] = gcrow[cl] + h_nlcl[cl] + 6 * (d - cl * P->nsub_cl);
    abase[d] = croff[d] + h_nlc[d];
  }
  std::vector<int> lcmap;
  d_lcmap.download(lcmap, st);
  std::vector<int64_t> lcrow(off_lc[N]), lcpos(off_lc[N]);
  for (int d = 0; d < N; d++) {
    const int cl = h_cl[d];
    for (int m = 0; m < h_nlc[d]; m++) {
      const int64_t e = off_lc[d] + m;
      lcrow[e] = lcmap[e] - h_la_off[cl] + gcrow[cl];
      lcpos[e] = croff[d] + m;
    }
  }
  b_obrow.upload(obrow, st);
  {  // b-block row of every b position
    std::vector<int> qrow(NB);
    for (int d = 0; d < N; d++)
      for (int p = 0; p < h_ns[d]; p++) qrow[off_b[d] + p] = (int)(obrow[d] + p);
    b_qrow.upload(qrow, st);
    // b-block rows of each dual row's + and - entries (kb_scatter)
    std::vector<int> ra, rb;
    d_ra.download(ra, st);
    d_rb.download(rb, st);
    std::vector<int> srow(2 * P->nlam);
    for (int64_t r = 0; r < P->nlam; r++) {
      srow[2 * r] = qrow[ra[r]];
      srow[2 * r + 1] = rb[r] >= 0 ? qrow[rb[r]] : -1;
    }
    b_srow.upload(srow, st);
  }
  b_orrow.upload(orrow, st);
  b_xoff.upload(xoff, st);
  b_roff.upload(roff, st);
  b_croff.upload(croff_el, st);
  b_crrow.upload(croff, st);
  b_cro.upload(cro, st);
  b_gcrow.upload(gcrow, st);
  b_gcoff.upload(gcoff, st);
  b_lcrow.upload(lcrow, st);
  b_lcpos.upload(lcpos, st);
  b_arow.upload(arow, st);
  b_abase_cr.upload(abase, st);
  b_v6off.upload(v6off, st);
  b_v6row.upload(v6row, st);
  b_crn.upload(crn, st);
  // work blocks of block context 0 (zero: padded rows and unused columns must stay finite)
  b_nparts = (int)((nl + 63) / 64);
  if (ctxs_.empty()) ctxs_.emplace_back(new BlockCtx);  // placeholder of the current context 0
  alloc_ctx_buffers();
  // CRQ = [C_i | R_b,i | Q_i | Yh_i]: n_s x (nlc + 18), column major, ld = ldb, with
  // Q_i an orthonormal basis of span(R_b,i) and, for A = K+_bb, Y = A Q, W = Q^T Y,
  // Yh = Y - Q W / 2, so that A = (I - QQ^T) A (I - QQ^T) + Q Yh^T + Yh Q^T.  The
  // projected part is ~100x smaller than A (the regularised inverse is large along
  // the rigid-body modes) and is the part stored in FP32.
  zero_alloc(b_CR, co, st);
  if (maxnlc > 0)
    HC(launch_copy_cols(st, N, maxnlc, d_ns.p, d_nlc.p, d_C.p, d_off_C.p, d_ldb.p, b_CR.p, b_cro.p, d_ldb.p, 1.0));
  {
    std::vector<double> Rb, Q(off_Rb[N]), Y;
    d_Rb.download(Rb, st);
    {  // per dual row, the rows of R_b of its two subdomains (the projector's G rows)
      std::vector<double> Gr((size_t)P->nlam * 12, 0.0);
      for (int d = 0; d < N; d++) {
        const Subdomain& sd = P->subs[d];
        for (size_t e = 0; e < sd.bf_row.size(); e++) {
          const int64_t r = sd.bf_row[e];
          const int p = sd.bpos[sd.bf_dof[e]];
          const int side = sd.bf_sign[e] > 0 ? 0 : 6;
          for (int m = 0; m < 6; m++) Gr[(size_t)r * 12 + side + m] = Rb[off_Rb[d] + p + (int64_t)m * h_ns[d]];
        }
      }
      b_Gr.upload(Gr, st);
      // per group of 64 dual rows, the subdomains whose h rows kb_gh stages in LDS
      // (first GH_SLOTS distinct ones in row order) and each row's two slots
      std::vector<int> ra(P->nlam, -1), rb(P->nlam, -1);
      for (int d = 0; d < N; d++) {
        const Subdomain& sd = P->subs[d];
        for (size_t e = 0; e < sd.bf_row.size(); e++) (sd.bf_sign[e] > 0 ? ra : rb)[sd.bf_row[e]] = d;
      }
      const int64_t ngh = (P->nlam + 63) / 64;
      std::vector<int> gsub((size_t)ngh * GH_SLOTS, -1), rslot(P->nlam, 0);
      for (int64_t g = 0; g < ngh; g++) {
        int* gs = gsub.data() + g * GH_SLOTS;
        int cnt = 0;
        auto slot = [&](int dd) {
          for (int q = 0; q < cnt; q++)
            if (gs[q] == dd) return q;
          if (cnt < GH_SLOTS) {
            gs[cnt] = dd;
            return cnt++;
          }
          return 255;
        };
        for (int64_t r = 64 * g; r < std::min<int64_t>(P->nlam, 64 * g + 64); r++) {
          const int sa = slot(ra[r]);
          const int sb = slot(rb[r] >= 0 ? rb[r] : ra[r]);
          rslot[r] = sa | (sb << 8);
        }
      }
      b_ghsub.upload(gsub, st);
      b_ghslot.upload(rslot, st);
    }
    for (int d = 0; d < N; d++) {  // modified Gram-Schmidt per subdomain
      const int n = h_ns[d];
      double* q = Q.data() + off_Rb[d];
      std::copy(Rb.begin() + off_Rb[d], Rb.begin() + off_Rb[d + 1], q);
      for (int a = 0; a < 6; a++) {
        for (int b = 0; b < a; b++) {
          double t = 0.0;
          for (int p = 0; p < n; p++) t += q[p + (int64_t)a * n] * q[p + (int64_t)b * n];
          for (int p = 0; p < n; p++) q[p + (int64_t)a * n] -= t * q[p + (int64_t)b * n];
        }
        double t = 0.0;
        for (int p = 0; p < n; p++) t += q[p + (int64_t)a * n] * q[p + (int64_t)a * n];
        if (!(t > 0.0)) throw std::runtime_error("interface kernel block is rank deficient");
        t = 1.0 / std::sqrt(t);
        for (int p = 0; p < n; p++) q[p + (int64_t)a * n] *= t;
      }
    }
    DVec<double> dQ, dY;
    dQ.upload(Q, st);
    dY.alloc(off_Rb[N]);
    HC(launch_kb_cols(st, N, maxns, 0, 6, KP, d_ns.p, b_obrow.p, dQ.p, d_off_Rb.p, d_ns.p, b_xb.p));
    HC(launch_tgemm_f64(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, d_Kt.p, d_off_P.p, nullptr, b_xb.p, b_xoff.p, b_yb.p,
                        b_xoff.p, nullptr, KP, 1.0, nullptr, 1, nullptr, nullptr, b_xob.p, n_xob));
    HC(launch_kb_cols(st, N, maxns, 1, 6, KP, d_ns.p, b_obrow.p, dY.p, d_off_Rb.p, d_ns.p, b_yb.p));
    dY.download(Y, st);
    for (int d = 0; d < N; d++) {
      const int n = h_ns[d];
      const double* q = Q.data() + off_Rb[d];
      double* y = Y.data() + off_Rb[d];
      double W[36];
      for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
          double t = 0.0;
          for (int p = 0; p < n; p++) t += q[p + (int64_t)a * n] * y[p + (int64_t)b * n];
          W[a * 6 + b] = t;
        }
      for (int b = 0; b < 6; b++)
        for (int a = 0; a < 6; a++) {
          const double w = 0.25 * (W[a * 6 + b] + W[b * 6 + a]);  // half of the symmetrised W
          for (int p = 0; p < n; p++) y[p + (int64_t)b * n] -= q[p + (int64_t)a * n] * w;
        }
    }
    dY.upload(Y, st);
    DVec<int64_t> dcr, dcq, dcy;
    dcr.upload(cro_rb, st);
    dcq.upload(cro_q, st);
    dcy.upload(cro_y, st);
    HC(launch_copy_cols(st, N, 6, d_ns.p, d_sixes.p, d_Rb.p, d_off_Rb.p, d_ns.p, b_CR.p, dcr.p, d_ldb.p, 1.0));
    HC(launch_copy_cols(st, N, 6, d_ns.p, d_sixes.p, dQ.p, d_off_Rb.p, d_ns.p, b_CR.p, dcq.p, d_ldb.p, 1.0));
    HC(launch_copy_cols(st, N, 6, d_ns.p, d_sixes.p, dY.p, d_off_Rb.p, d_ns.p, b_CR.p, dcy.p, d_ldb.p, 1.0));
    // the same block as general 64x64 tiles for the tile GEMMs
    std::vector<int64_t> crto(N);
    int64_t ct = 0;
    int maxt = 0;
    for (int d = 0; d < N; d++) {
      crto[d] = ct;
      const int t = ((h_ns[d] + 63) / 64) * ((crn[d] + 63) / 64);
      ct += (int64_t)t * 4096;
      maxt = std::max(maxt, t);
    }
    b_crto.upload(crto, st);
    b_CRt.alloc(std::max<int64_t>(ct, 1));
    HC(launch_pack_f64(st, N, maxt, 1, d_ns.p, b_crn.p, b_CR.p, b_cro.p, d_ldb.p, b_CRt.p, b_crto.p));
    HC(hipMemsetAsync(b_xb.p, 0, b_xb.n * sizeof(double), st));
    HC(hipMemsetAsync(b_yb.p, 0, b_yb.n * sizeof(double), st));
    HC(hipStreamSynchronize(st));
  }
  // FP32 tiles: deflated K+_bb, plain S_i
  b_Kt32.alloc(off_P[N]);
  b_St32.alloc(off_P[N]);
  HC(launch_kb_deflate_f32(st, N, maxtiles, d_ns.p, d_Kt.p, d_off_P.p, b_CR.p, b_cro.p, d_ldb.p, d_nlc.p, b_Kt32.p));
  HC(launch_kb_to_f32(st, off_P[N], d_St.p, b_St32.p));
  if (gemm16) {  // in place: FP32 tiles -> hi and lo FP16 planes with a power-of-two scale per subdomain
    b_sK.alloc(N);
    b_sS.alloc(N);
    HC(launch_split16(st, N, maxtiles, d_ns.p, d_off_P.p, b_Kt32.p, b_sK.p));
    HC(launch_split16(st, N, maxtiles, d_ns.p, d_off_P.p, b_St32.p, b_sS.p));
  }
  HC(hipStreamSynchronize(st));
  autotune();
  blk_ready = true;
}

// the buffers of the current block context
void Engine::alloc_ctx_buffers() {
  const int N = P->nsub;
  const int64_t nl = P->nlam, np = off_n[N], ng = 3 * (int64_t)P->nPoints;
  zero_alloc(b_xb, NBP * KP, st);
  zero_alloc(b_yb, NBP * KP, st);
  zero_alloc(b_xb32, NBP * KP, st);
  zero_alloc(b_yb32, NBP * KP, st);
  zero_alloc(b_sx, (int64_t)N * KP, st);
  if (gemm16) zero_alloc(b_p16, std::max<int64_t>(p16_n, 1), st);
  zero_alloc(b_cr, CRP * KP, st);
  zero_alloc(b_coef, CRP * KP, st);
  zero_alloc(b_kpart, std::max<int64_t>((int64_t)N * 8, (int64_t)P->ncl * maxTcl * sAc) * 64 * KP, st);
  zero_alloc(b_v6, (int64_t)64 * N * KP, st);
  zero_alloc(b_gc, GCP * KP, st);
  zero_alloc(b_la, GCP * KP, st);
  zero_alloc(b_h, (int64_t)6 * N * KMAX, st);
  zero_alloc(b_hg, (int64_t)6 * N * KMAX, st);
  zero_alloc(b_hgf, (int64_t)6 * N * KMAX, st);
  for (DVec<double>* v : {&b_lam, &b_g, &b_w, &b_Fw, &b_Pg, &b_z, &b_Pz, &b_d, &b_g64}) zero_alloc(*v, nl * KMAX, st);
  zero_alloc(b_part, (int64_t)b_nparts * KMAX, st);
  zero_alloc(b_gt, (int64_t)6 * N * KP, st);
  zero_alloc(b_gtp, (int64_t)6 * std::max(n_gchunk, 1) * KP, st);
  zero_alloc(b_Vr, NRP * KP, st);
  zero_alloc(b_Xr, NRP * KP, st);
  zero_alloc(b_Tr, NRP * KP, st);
  zero_alloc(b_V, np * KMAX, st);
  zero_alloc(b_V2, np * KMAX, st);
  zero_alloc(b_X, np * KMAX, st);
  zero_alloc(b_U, ng * KMAX, st);
  b_sc.alloc(1);
}

void Engine::swap_ctx(BlockCtx& c) {
  std::swap(st, c.st);
  std::swap(gx_blk, c.gx);
  std::swap(gx_blk_key, c.gx_key);
  std::swap(blk_k, c.k);
  std::swap(blk_prec, c.prec);
  b_xb32.swap(c.xb32);
  b_yb32.swap(c.yb32);
  b_sx.swap(c.sx);
  b_p16.swap(c.p16);
  b_xb.swap(c.xb); b_yb.swap(c.yb); b_cr.swap(c.cr); b_coef.swap(c.coef); b_kpart.swap(c.kpart); b_v6.swap(c.v6);
  b_gc.swap(c.gc); b_la.swap(c.la); b_h.swap(c.h); b_hg.swap(c.hg); b_hgf.swap(c.hgf); b_lam.swap(c.lam);
  b_g.swap(c.g); b_w.swap(c.w); b_Fw.swap(c.Fw); b_Pg.swap(c.Pg); b_z.swap(c.z); b_Pz.swap(c.Pz); b_d.swap(c.d);
  b_g64.swap(c.g64); b_part.swap(c.part); b_gt.swap(c.gt); b_gtp.swap(c.gtp); b_Vr.swap(c.Vr); b_Xr.swap(c.Xr); b_Tr.swap(c.Tr);
  b_V.swap(c.V); b_V2.swap(c.V2); b_X.swap(c.X); b_U.swap(c.U);
  b_sc.swap(c.sc);
}

void Engine::use_ctx(int c) {
  if (c == cur_ctx_) return;
  if (c < 0 || c >= (int)ctxs_.size()) throw std::runtime_error("no such block context");
  swap_ctx(*ctxs_[cur_ctx_]);  // the current context into its slot, its placeholder into the engine
  swap_ctx(*ctxs_[c]);         // context c into the engine
  cur_ctx_ = c;
}

int Engine::blk_contexts(int n) {
  if (n < 1 || n > 8) throw std::runtime_error("block contexts must be 1..8");
  use_ctx(0);
  block_prepare();
  while ((int)ctxs_.size() < n) {
    std::unique_ptr<BlockCtx> c(new BlockCtx);
    HC(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    ctxs_.push_back(std::move(c));
    use_ctx((int)ctxs_.size() - 1);
    alloc_ctx_buffers();
    HC(hipStreamSynchronize(st));
    use_ctx(0);
  }
  return (int)ctxs_.size();
}

// Autotuner of the block GEMM tile shape: every candidate is timed on this mesh's
// own K+_bb tiles over a 64-column block with HIP events, and the fastest is used
// for K+_bb and S_i from then on.  Candidates 1, 2: FP32 format 64 or 128 output rows
// per workgroup; split-FP16 format K chunks of 32 or 64 per LDS stage.
void Engine::autotune() {
  const int N = P->nsub;
  Events ev(2);
  hipEvent_t a = ev[0], b = ev[1];
  for (int rt = 1; rt <= 2; rt++) {
    auto launch = [&] {
      if (gemm16)
        HC(launch_tgemm_h16(st, N, maxTb, d_ns.p, b_Kt32.p, d_off_P.p, b_sK.p, b_xb32.p, b_xoff.p, b_sx.p, KP, 1.0f,
                            nullptr, b_yb32.p, b_xob.p, n_xob, 32 * rt));
      else
        HC(launch_tgemm_f32(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, b_Kt32.p, d_off_P.p, nullptr, b_xb32.p, b_xoff.p,
                            nullptr, b_xoff.p, nullptr, KP, 1.0f, nullptr, rt, b_yb32.p, b_xob.p, n_xob));
    };
    launch();
    HC(hipEventRecord(a, st));
    for (int r = 0; r < 3; r++) launch();
    HC(hipEventRecord(b, st));
    HC(hipEventSynchronize(b));
    float ms = 0;
    HC(hipEventElapsedTime(&ms, a, b));
    tune_ms32[rt - 1] = ms / 3;
  }
  tune_rt32 = gemm_rows_forced ? gemm_rows_forced : (tune_ms32[1] < tune_ms32[0] ? 2 : 1);
  if (gemm16) {  // the band form, timed the same way (chosen by the gemm_kernel setting, not by time)
    auto launch = [&] {
      HC(launch_tgemm_h16b(st, N, maxTb, d_ns.p, b_Kt32.p, d_off_P.p, b_sK.p, b_xb32.p, b_xoff.p, b_sx.p, KP, 1.0f,
                           nullptr, b_yb32.p, b_xob.p, n_xob, 1, b_p16.p, d_p16off.p));
    };
    launch();
    HC(hipEventRecord(a, st));
    for (int r = 0; r < 3; r++) launch();
    HC(hipEventRecord(b, st));
    HC(hipEventSynchronize(b));
    float ms = 0;
    HC(hipEventElapsedTime(&ms, a, b));
    tune_ms32[2] = ms / 3;
  }
}

size_t Engine::block_bytes() const {
  if (!blk_ready) return 0;
  size_t s = 0;
  s += (b_Kt32.n + b_St32.n + b_xb32.n + b_yb32.n + b_sK.n + b_sS.n + b_sx.n + b_p16.n) * sizeof(float);
  for (const DVec<double>* v : {&b_CR, &b_CRt, &b_kpart, &b_Gr, &b_gt, &b_gtp, &b_xb, &b_yb, &b_cr, &b_coef, &b_v6, &b_gc, &b_la, &b_h, &b_hg, &b_hgf, &b_lam, &b_g, &b_w,
                                &b_Fw, &b_Pg, &b_z, &b_Pz, &b_d, &b_g64, &b_part, &b_Vr, &b_Xr, &b_Tr, &b_V, &b_V2, &b_X, &b_U})
    s += v->n * sizeof(double);
  s += (b_ghsub.n + b_ghslot.n + b_srow.n) * sizeof(int);
  return s;
}

// out = F win (- sub), with per-column partial dots against dotw; every launch is
// skipped on the device while *gate == 0 (gate may be null)
void Engine::bF(int k, int prec, const double* win, double* out, const double* sub, const double* dotw,
                const int* gate) {
  const int N = P->nsub, NC = P->ncl;
  HC(launch_kb_gather(st, NB, k, KP, b_qrow.p, d_gptr.p, d_gidx.p, win, nullptr, b_xb.p,
                      prec == 32 ? b_xb32.p : nullptr, gate));
  if (prec == 32) {  // FP32 result into b_yb32; the FP64 correction below adds it into b_yb
    gemm_lo(0, gate);
  } else {
    HC(launch_tgemm_f64(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, d_Kt.p, d_off_P.p, nullptr, b_xb.p, b_xoff.p, b_yb.p,
                        b_xoff.p, nullptr, KP, 1.0, gate, 1, nullptr, nullptr, b_xob.p, n_xob));
  }
  // HTFETI correction (Cluster.cpp:1129-1186): [c; a] = [C_i | R_b,i]^T x_b,
  // [lam_c; alpha] = Ac^-1 [c; -a], y_b += R_b alpha - C lam_c; in FP32 also the
  // rank-12 FP64 part of K+_bb, Q (Yh^T x) + Yh (Q^T x)
  crq_products(gate);
  HC(launch_kb_cluster_gc(st, NC, max_cl_len, k, KP, d_gc_src.p, d_gc_off.p, d_nlcl.p, d_nrcl.p, d_d0.p, b_cr.p,
                          b_lcpos.p, b_cr.p, b_abase_cr.p, b_gc.p, b_gcrow.p, gate));
  HC(launch_tgemm_f64(st, NC, maxTcl, 2, 0, d_nlen.p, d_nlen.p, d_Ainv.p, d_Aoff.p, d_nlen.p, b_gc.p, b_gcoff.p,
                      b_la.p, b_gcoff.p, nullptr, KP, 1.0, gate, sAc, b_kpart.p));
  b_corr(k, 1, prec == 32 ? 1 : 0, gate, prec == 32 ? b_yb32.p : nullptr);
  HC(launch_kb_scatter(st, P->nlam, k, KP, b_srow.p, b_yb.p, nullptr, sub, out,
                       dotw, b_part.p, gate));
}

void Engine::gemm_lo(int which, const int* gate) {
  const int N = P->nsub;
  const float* A = which == 0 ? b_Kt32.p : b_St32.p;
  if (gemm_probe && which == 0 && gemm_probe_n + 2 <= gemm_probe->e.size()) {
    // the column scales (split FP16), then the GEMM alone between two events: the
    // band kernel and its reduction, the row form or the FP32 tile GEMM
    if (gemm16) HC(launch_kb_colscale(st, N, d_ns.p, b_xoff.p, KP, b_xb32.p, b_sx.p, gate));
    HC(hipEventRecord(gemm_probe->e[gemm_probe_n++], st));
    if (gemm16 && gemm_band)
      HC(launch_tgemm_h16b(st, N, maxTb, d_ns.p, A, d_off_P.p, b_sK.p, b_xb32.p, b_xoff.p, b_sx.p, KP, 1.0f, gate,
                           b_yb32.p, b_xob.p, n_xob, 0, b_p16.p, d_p16off.p));
    else if (gemm16)
      HC(launch_tgemm_h16(st, N, maxTb, d_ns.p, A, d_off_P.p, b_sK.p, b_xb32.p, b_xoff.p, b_sx.p, KP, 1.0f, gate,
                          b_yb32.p, b_xob.p, n_xob, 32 * tune_rt32, 0, gemm_lds_pad, gemm_exact, gemm_pf));
    else
      HC(launch_tgemm_f32(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, A, d_off_P.p, nullptr, b_xb32.p, b_xoff.p, nullptr,
                          b_xoff.p, nullptr, KP, 1.0f, gate, tune_rt32, b_yb32.p, b_xob.p, n_xob));
    HC(hipEventRecord(gemm_probe->e[gemm_probe_n++], st));
  } else if (gemm16 && gemm_band)
    HC(launch_tgemm_h16b(st, N, maxTb, d_ns.p, A, d_off_P.p, which == 0 ? b_sK.p : b_sS.p, b_xb32.p, b_xoff.p, b_sx.p,
                         KP, 1.0f, gate, b_yb32.p, b_xob.p, n_xob, 1, b_p16.p, d_p16off.p));
  else if (gemm16)
    HC(launch_tgemm_h16(st, N, maxTb, d_ns.p, A, d_off_P.p, which == 0 ? b_sK.p : b_sS.p, b_xb32.p, b_xoff.p, b_sx.p,
                        KP, 1.0f, gate, b_yb32.p, b_xob.p, n_xob, 32 * tune_rt32, 1, gemm_lds_pad, gemm_exact, gemm_pf));
  else
    HC(launch_tgemm_f32(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, A, d_off_P.p, nullptr, b_xb32.p, b_xoff.p, nullptr,
                        b_xoff.p, nullptr, KP, 1.0f, gate, tune_rt32, b_yb32.p, b_xob.p, n_xob));
}

// y_b += R_b alpha (with_R) - C lam_c (+ Q (Yh^T x) + Yh (Q^T x) if defl) as one
// FP64 tile GEMM of CRQ with the coefficient block gathered from la and cr
void Engine::b_corr(int k, int with_R, int defl, const int* gate, const float* yf) {
  const int N = P->nsub;
  HC(launch_kb_coef(st, N, maxnlc + 18, k, KP, d_nlc.p, d_off_lc.p, b_lcrow.p, b_arow.p, b_la.p, with_R, defl, b_cr.p,
                    b_crrow.p, b_coef.p, gate));
  if (maxTcr == 1)  // at most 64 correction columns: coefficients staged once per row block
    HC(launch_corr_f64(st, N, maxTb, d_ns.p, b_crn.p, b_CRt.p, b_crto.p, b_coef.p, b_croff.p, b_yb.p, b_xoff.p,
                       yf ? nullptr : b_yb.p, yf, KP, gate, b_xob.p, n_xob));
  else
    HC(launch_tgemm_f64(st, N, maxTb, 1, 0, d_ns.p, b_crn.p, b_CRt.p, b_crto.p, nullptr, b_coef.p, b_croff.p, b_yb.p,
                        b_xoff.p, yf ? nullptr : b_yb.p, KP, 1.0, gate, 1, nullptr, yf, b_xob.p, n_xob));
}

// cr = CRQ^T x_b per subdomain (c = C^T x, a = R_b^T x, Q^T x, Yh^T x), from the
// CRQ tiles; split over K when the product is a single 64-row block
void Engine::crq_products(const int* gate) {
  const int N = P->nsub;
  const int S = maxTcr == 1 ? 8 : 1;
  HC(launch_tgemm_f64(st, N, maxTcr, 1, 1, d_ns.p, b_crn.p, b_CRt.p, b_crto.p, nullptr, b_xb.p, b_xoff.p, b_cr.p,
                      b_croff.p, nullptr, KP, 1.0, gate, S, b_kpart.p, nullptr, b_xob.p, n_xob));
}

// hbuf = H (G^T G)^-1 G^T x per subdomain, with G^T x = R_b^T (Bf^T x) taken
// straight from the dual block (Cluster.cpp:1314-1329)
void Engine::b_gtx_coarse(int k, const double* x, double* hbuf, const int* gate) {
  const int N = P->nsub;
  HC(launch_kb_gtx(st, N, n_gchunk, k, KP, d_ns.p, b_gcsub.p, b_gcp0.p, b_gcbeg.p, d_off_b.p, d_gptr.p, d_gidx.p,
                   d_gpos.p, d_Rb.p, d_off_Rb.p, x, b_gtp.p, b_gt.p, gate));
  HC(launch_kb_proj_coarse(st, N, 6 * P->ncl, k, KP, b_gt.p, nullptr, d_H.p, d_cl.p, d_invGtG.p, hbuf, nullptr, 1.0,
                           gate));
}

// out = P x = x - G (G^T G)^-1 G^T x (Cluster.cpp:1314-1329); hbuf = H alpha per subdomain
void Engine::bP(int k, const double* x, double* out, double* hbuf, const double* dotv, const int* gate) {
  b_gtx_coarse(k, x, hbuf, gate);
  HC(launch_kb_gh(st, P->nlam, k, d_rsa.p, d_rsb.p, b_ghsub.p, b_ghslot.p, b_Gr.p, hbuf, x, out, -1.0, dotv, b_part.p, gate));
}

// out = M P x (Cluster.cpp:1043-1112 applied to the projection of x); P x goes
// through the dual temporary b_Pg
void Engine::bPM(int k, int prec, const double* x, double* out, double* hbuf, const int* gate) {
  const int N = P->nsub;
  b_gtx_coarse(k, x, hbuf, gate);
  HC(launch_kb_gh(st, P->nlam, k, d_rsa.p, d_rsb.p, b_ghsub.p, b_ghslot.p, b_Gr.p, hbuf, x, b_Pg.p, -1.0, nullptr, nullptr, gate));
  HC(launch_kb_gather(st, NB, k, KP, b_qrow.p, d_gptr.p, d_gidx.p, b_Pg.p, d_wgt.p,
                      prec == 64 ? b_xb.p : nullptr, prec == 32 ? b_xb32.p : nullptr, gate));
  if (prec == 32) {
    gemm_lo(1, gate);
  } else {
    HC(launch_tgemm_f64(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, d_St.p, d_off_P.p, nullptr, b_xb.p, b_xoff.p, b_yb.p,
                        b_xoff.p, nullptr, KP, 1.0, gate, 1, nullptr, nullptr, b_xob.p, n_xob));
  }
  HC(launch_kb_scatter(st, P->nlam, k, KP, b_srow.p, b_yb.p, d_wgt.p, nullptr,
                       out, nullptr, nullptr, gate, prec == 32 ? b_yb32.p : nullptr));
}

// out = W Bf S Bf^T W x (Cluster.cpp:1043-1112)
void Engine::bM(int k, int prec, const double* x, double* out, const int* gate) {
  const int N = P->nsub;
  HC(launch_kb_gather(st, NB, k, KP, b_qrow.p, d_gptr.p, d_gidx.p, x, d_wgt.p,
                      prec == 64 ? b_xb.p : nullptr, prec == 32 ? b_xb32.p : nullptr, gate));
  if (prec == 32) {
    gemm_lo(1, gate);
  } else {
    HC(launch_tgemm_f64(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, d_St.p, d_off_P.p, nullptr, b_xb.p, b_xoff.p, b_yb.p,
                        b_xoff.p, nullptr, KP, 1.0, gate, 1, nullptr, nullptr, b_xob.p, n_xob));
  }
  HC(launch_kb_scatter(st, P->nlam, k, KP, b_srow.p, b_yb.p, d_wgt.p, nullptr,
                       out, nullptr, nullptr, gate, prec == 32 ? b_yb32.p : nullptr));
}

// X <- K_rr^-1 X = W^T (W X) on an interior block (rows orrow, ld KP), as two
// lower-triangular tile GEMMs with the kept W = L_rr^-1
void Engine::b_rr_solve(double* X) {
  const int N = P->nsub;
  HC(launch_tgemm_f64(st, N, maxTr, 3, 0, d_nr.p, d_nr.p, d_Lt.p, d_off_L.p, nullptr, X, b_roff.p, b_Tr.p, b_roff.p,
                      nullptr, KP, 1.0, nullptr, 1, nullptr, nullptr, b_xor.p, n_xor));
  HC(launch_tgemm_f64(st, N, maxTr, 3, 1, d_nr.p, d_nr.p, d_Lt.p, d_off_L.p, nullptr, b_Tr.p, b_roff.p, X, b_roff.p,
                      nullptr, KP, 1.0, nullptr, 1, nullptr, nullptr, b_xor.p, n_xor));
}

// From the local block V (FP64, column major): b_Vr = K_rr^-1 v_r, b_xb = t_b = v_b - K_br K_rr^-1 v_r,
// b_yb = (K^+ (v - Bc^T lam_c))_b (+ R_b alpha_c if with_R), b_la = [lam_c; alpha_c],
// b_v6 = R^T v (Cluster.cpp:967-1040).  With interior_too, b_X = K^+(v - Bc^T lam_c)
// in local order (packed, column major).
// interior_too is the recovery, whose V = f - Bf^T lam differs from the loads f only
// at B-touched DOFs: its v_r is the loads' f_r, and b_Vr = K_rr^-1 f_r as the last
// initialisation left it (every join initialises all slots from their loads) is used
// again instead of a second interior solve.
void Engine::b_kplus(int k, bool interior_too, double* V) {
  const int N = P->nsub, NC = P->ncl;
  const int64_t np = off_n[N];
  HC(launch_kb_permute(st, N, maxn, 0, k, KP, np, d_n.p, d_nr.p, d_perm1.p, d_off_n.p, V,
                       interior_too ? nullptr : b_Vr.p, b_orrow.p, b_xb.p, b_obrow.p));
  // b_Vr = K_rr^-1 v_r = W^T (W v_r) ; t_b = v_b - K_br K_rr^-1 v_r
  if (maxTr > 0) {
    if (!interior_too) b_rr_solve(b_Vr.p);
    HC(launch_csr_apply(st, 0, NB, 0, k, d_brp.p, d_brc.p, d_brv.p, d_bsub.p, d_bloc.p, b_Vr.p, b_roff.p, KP, 1,
                        b_xb.p, b_xoff.p, KP, 1, nullptr, -1.0, 1.0));
  }
  HC(launch_tgemm_f64(st, N, maxTb, 0, 0, d_ns.p, d_ns.p, d_Kt.p, d_off_P.p, nullptr, b_xb.p, b_xoff.p, b_yb.p,
                      b_xoff.p, nullptr, KP, 1.0, nullptr, 1, nullptr, nullptr, b_xob.p, n_xob));
  crq_products(nullptr);
  HC(launch_kb_rtf(st, N, k, KP, np, d_n.p, d_off_n.p, d_R.p, d_off_R.p, V, b_v6.p));
  HC(launch_kb_cluster_gc(st, NC, max_cl_len, k, KP, d_gc_src.p, d_gc_off.p, d_nlcl.p, d_nrcl.p, d_d0.p, b_cr.p,
                          b_lcpos.p, b_v6.p, b_v6row.p, b_gc.p, b_gcrow.p));
  HC(launch_tgemm_f64(st, NC, maxTcl, 2, 0, d_nlen.p, d_nlen.p, d_Ainv.p, d_Aoff.p, d_nlen.p, b_gc.p, b_gcoff.p,
                      b_la.p, b_gcoff.p, nullptr, KP, 1.0, nullptr, sAc, b_kpart.p));
  b_corr(k, interior_too ? 0 : 1, 0, nullptr);
  if (!interior_too) return;
  // x_r = K_rr^-1 v_r - K_rr^-1 K_rb y_b
  if (maxTr > 0) {
    HC(launch_csr_apply(st, 0, NR, 0, k, d_rbp.p, d_rbc.p, d_rbv.p, d_rsub.p, d_rloc.p, b_yb.p, b_xoff.p, KP, 1,
                        b_Xr.p, b_roff.p, KP, 1, nullptr, 1.0, 0.0));
    b_rr_solve(b_Xr.p);
    HC(launch_rr_axpby(st, N, maxnr, k, d_nr.p, b_Vr.p, b_roff.p, b_Xr.p, b_roff.p, KP, 1, 1.0, -1.0));
  }
  HC(launch_kb_permute(st, N, maxn, 1, k, KP, np, d_n.p, d_nr.p, d_perm1.p, d_off_n.p, b_X.p, b_Xr.p, b_orrow.p,
                       b_yb.p, b_obrow.p));
}

void Engine::apply_block(int op, int k, int prec, const double* x_h, double* y_h) {
  if (k < 1 || k > KMAX) throw std::runtime_error("block size must be 1..64");
  if (prec != 32 && prec != 64) throw std::runtime_error("precision must be 32 or 64");
  if (op < 0 || op > 2) throw std::runtime_error("operator must be 0 (F), 1 (M) or 2 (P)");
  block_prepare();
  const int64_t nl = P->nlam;
  std::vector<double> t((size_t)nl * k);
  for (int j = 0; j < k; j++)
    for (int64_t r = 0; r < nl; r++) t[(size_t)r * k + j] = x_h[(size_t)j * nl + r];
  HC(hipMemcpyAsync(b_w.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, st));
  if (op == 0) bF(k, prec, b_w.p, b_Fw.p, nullptr, nullptr);
  else if (op == 1) bM(k, prec, b_w.p, b_Fw.p);
  else bP(k, b_w.p, b_Fw.p, b_h.p, nullptr);
  HC(hipMemcpyAsync(t.data(), b_Fw.p, t.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
  for (int j = 0; j < k; j++)
    for (int64_t r = 0; r < nl; r++) y_h[(size_t)j * nl + r] = t[(size_t)r * k + j];
}

double Engine::time_gemm_in_solve(int k, int* nlaunch) {
  k = std::max(1, std::min(k, KMAX));
  int64_t np = 0;
  for (const Subdomain& sd : P->subs) np += sd.n;
  std::vector<double> f((size_t)k * np);
  traction_loads(*P, 20240611ull, 0, k, f.data());
  std::vector<double> u((size_t)k * 3 * P->nPoints);
  Events ev(2 * 4096);
  const bool g = use_graph;
  use_graph = false;  // events between the launches need the eager launch sequence
  gemm_probe = &ev;
  gemm_probe_n = 0;
  try {
    solve_block(1e-8, 1000, k, 32, f.data(), u.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
  } catch (...) {
    gemm_probe = nullptr;
    use_graph = g;
    throw;
  }
  gemm_probe = nullptr;
  use_graph = g;
  HC(hipStreamSynchronize(st));
  // launches skipped through their device gate (no running column) take microseconds:
  // the mean over the launches longer than a quarter of the longest one
  std::vector<float> ms(gemm_probe_n / 2);
  float mx = 0;
  for (size_t i = 0; i < ms.size(); i++) {
    HC(hipEventElapsedTime(&ms[i], ev[2 * i], ev[2 * i + 1]));
    mx = std::max(mx, ms[i]);
  }
  double sum = 0;
  int n = 0;
  for (float x : ms)
    if (x > 0.25f * mx) sum += x, n++;
  if (nlaunch) *nlaunch = n;
  return n ? sum / n : 0.0;
}

double Engine::time_block_gemm(int which, int reps, double* flops) {
  block_prepare();
  const int N = P->nsub;
  const int64_t ncopy = b_xb.n & ~(int64_t)1;
  // the matrix-core GEMM alone: in the split-FP16 format the column scales of X are
  // computed by the untimed first call
  auto launch = [&] {
    if (which == 2)
      HC(launch_kb_copy(st, ncopy, b_xb.p, b_yb.p));
    else if (gemm16 && gemm_band)  // the band kernel and its partial-sum reduction
      HC(launch_tgemm_h16b(st, N, maxTb, d_ns.p, which == 0 ? b_Kt32.p : b_St32.p, d_off_P.p,
                           which == 0 ? b_sK.p : b_sS.p, b_xb32.p, b_xoff.p, b_sx.p, KP, 1.0f, nullptr, b_yb32.p,
                           b_xob.p, n_xob, 0, b_p16.p, d_p16off.p));
    else if (gemm16)
      HC(launch_tgemm_h16(st, N, maxTb, d_ns.p, which == 0 ? b_Kt32.p : b_St32.p, d_off_P.p,
                          which == 0 ? b_sK.p : b_sS.p, b_xb32.p, b_xoff.p, b_sx.p, KP, 1.0f, nullptr, b_yb32.p,
                          b_xob.p, n_xob, 32 * tune_rt32, 0, gemm_lds_pad, gemm_exact, gemm_pf));
    else
      gemm_lo(which, nullptr);
  };
  if (which != 2) gemm_lo(which, nullptr);
  launch();
  Events ev(2);
  HC(hipEventRecord(ev[0], st));
  for (int i = 0; i < reps; i++) launch();
  HC(hipEventRecord(ev[1], st));
  HC(hipEventSynchronize(ev[1]));
  float ms = 0;
  HC(hipEventElapsedTime(&ms, ev[0], ev[1]));
  if (flops) {
    double f = 0;
    for (int d = 0; d < N; d++) f += 2.0 * KP * (double)h_ns[d] * h_ns[d];
    *flops = which == 2 ? 8.0 * (double)ncopy : f;  // the copy: bytes read (= bytes written)
  }
  return ms / reps;
}

// ------------------------------------------------------------ slot lifecycle
// The block is a set of k slots (columns).  A load case joins a free slot
// (blk_join: its d, lam0, g0, w0 and scalars are set up while the other columns
// keep their state), all running slots advance together (blk_sweeps), and a slot
// whose case has stopped is recovered and freed (blk_finish).  Joins and finishes
// happen between sweeps, so a case's arithmetic does not depend on the others.
void Engine::blk_begin(int k, double eps, int max_iter, int prec) {
  if (k < 1 || k > KMAX) throw std::runtime_error("block size must be 1..64");
  if (prec != 32 && prec != 64) throw std::runtime_error("precision must be 32 or 64");
  block_prepare();
  blk_k = k;
  blk_prec = prec;
  BScalars s0;
  std::memset(&s0, 0, sizeof(s0));
  s0.eps = eps;
  s0.max_iter = max_iter;
  s0.cert = prec == 32 ? 1 : 0;
  // mixed precision: FP64 refinement point when the recursive residual has fallen
  // refine_tol below the last true one (the FP32 operators are accurate to ~1e-7)
  s0.rtol = prec == 32 ? optd(P->o, "refine_tol") : 0.0;
  for (int j = 0; j < KMAX; j++) s0.done[j] = 4;
  HC(hipMemcpyAsync(b_sc.p, &s0, sizeof(BScalars), hipMemcpyHostToDevice, st));
}

void Engine::blk_set_status(const std::vector<int>& cols, int status) {
  static const int codes[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  for (int j : cols) {
    if (j < 0 || j >= blk_k) throw std::runtime_error("slot out of range");
    HC(hipMemcpyAsync(&b_sc.p->done[j], &codes[status], sizeof(int), hipMemcpyHostToDevice, st));
  }
}

void Engine::blk_join(const std::vector<int>& cols, const double* f_h) {
  blk_load(cols, f_h);
  blk_init(cols);
}

void Engine::blk_load(const std::vector<int>& cols, const double* f_h) {
  const int64_t np = off_n[P->nsub];
  for (size_t c = 0; c < cols.size(); c++) {
    double* dst = b_V.p + (size_t)cols[c] * np;
    if (f_h)
      HC(hipMemcpyAsync(dst, f_h + c * (size_t)np, np * sizeof(double), hipMemcpyHostToDevice, st));
    else
      HC(hipMemcpyAsync(dst, d_f.p, np * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
}

void Engine::blk_load_traction(const std::vector<int>& cols, uint64_t seed, const std::vector<int>& cases) {
  if (cols.size() != cases.size()) throw std::runtime_error("one load case per slot");
  const int NY = P->NE[1], NZ = P->NE[2], NX = P->NE[0];
  const int ny1 = NY + 1, nz1 = NZ + 1;
  const int64_t np = off_n[P->nsub], field = 3 * (int64_t)ny1 * nz1;
  if (n_face < 0) {  // the subdomains on the x = 1 face, once
    std::vector<int> meta;
    std::vector<int64_t> foff;
    for (int d = 0; d < P->nsub; d++) {
      const Subdomain& s = P->subs[d];
      if (s.e0[0] + s.ne[0] != NX) continue;
      meta.insert(meta.end(), {s.e0[1], s.ne[1], s.e0[2], s.ne[2], s.nn[0], s.nn[0] * s.nn[1]});
      foff.push_back(off_n[d]);
    }
    n_face = (int)foff.size();
    if (meta.empty()) meta.assign(6, 0), foff.assign(1, 0);
    d_face_meta.upload(meta, st);
    d_face_off.upload(foff, st);
  }
  if (d_trac.n < (size_t)(field * KMAX)) d_trac.alloc(field * KMAX);
  std::vector<double> t((size_t)field * cols.size());
  for (size_t c = 0; c < cols.size(); c++) traction_face(*P, seed, cases[c], t.data() + c * field);
  HC(hipMemcpyAsync(d_trac.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, st));
  for (size_t c = 0; c < cols.size(); c++) {
    if (cols[c] < 0 || cols[c] >= blk_k) throw std::runtime_error("slot out of range");
    double* dst = b_V.p + (size_t)cols[c] * np;
    HC(hipMemsetAsync(dst, 0, np * sizeof(double), st));
    HC(launch_traction_lump(st, n_face, d_face_meta.p, d_face_off.p, ny1, nz1, 1.0 / NY, 1.0 / NZ,
                            d_trac.p + c * field, dst));
  }
  HC(hipStreamSynchronize(st));  // the host field buffer is released on return
}

void Engine::blk_copy_u(int j, double* dst) {
  if (j < 0 || j >= blk_k) throw std::runtime_error("slot out of range");
  const int64_t ng = 3 * (int64_t)P->nPoints;
  HC(hipMemcpyAsync(dst, b_U.p + (size_t)j * ng, ng * sizeof(double), hipMemcpyDeviceToDevice, st));
}

void Engine::blk_init(const std::vector<int>& cols) {
  const int N = P->nsub, k = blk_k, prec = blk_prec;
  const int64_t nl = P->nlam;
  blk_set_status(cols, 5);
  // d = Bf K_c^+ f ; e = R^T f ; lam0 = G (G^T G)^-1 e (Cluster.cpp:1385-1394); computed
  // for every column into scratch, copied into the state of the joining ones
  b_kplus(k, false, b_V.p);
  HC(launch_kb_scatter(st, nl, k, KP, b_srow.p, b_yb.p, nullptr, nullptr, b_Fw.p,
                       nullptr, nullptr));
  HC(launch_kb_update(st, nl, k, 3, b_d.p, nullptr, b_Fw.p, nullptr, b_sc.p));
  HC(launch_kb_proj_coarse(st, N, 6 * P->ncl, k, KP, b_v6.p, b_v6row.p, d_H.p, d_cl.p, d_invGtG.p, b_h.p, nullptr,
                           1.0));
  HC(launch_kb_gh(st, nl, k, d_rsa.p, d_rsb.p, b_ghsub.p, b_ghslot.p, b_Gr.p, b_h.p, nullptr, b_Pg.p,
                  1.0, nullptr, nullptr));
  HC(launch_kb_update(st, nl, k, 3, b_lam.p, nullptr, b_Pg.p, nullptr, b_sc.p));
  // g0 = F lam0 - d in FP64 ; w0 = P M P g0
  bF(k, 64, b_lam.p, b_g64.p, b_d.p, nullptr);
  HC(launch_kb_update(st, nl, k, 3, b_g.p, nullptr, b_g64.p, nullptr, b_sc.p));
  bPM(k, prec, b_g64.p, b_z.p, b_hg.p);
  bP(k, b_z.p, b_Pz.p, b_h.p, b_g64.p);
  HC(launch_kb_update(st, nl, k, 3, b_w.p, nullptr, b_Pz.p, nullptr, b_sc.p));
  HC(launch_kb_reduce(st, b_nparts, k, b_part.p, b_sc.p, 2));
}

void Engine::blk_sweeps(int n) {
  const int k = blk_k, prec = blk_prec;
  const int64_t key = ((int64_t)n * 128 + k) * 2 + (prec == 32 ? 1 : 0);
  if (gx_blk && gx_blk_key != key) {
    (void)hipGraphExecDestroy(gx_blk);
    gx_blk = nullptr;
  }
  gx_blk_key = key;
  run_graph(&gx_blk, [&] { blk_enqueue_sweeps(n); });
}

void Engine::blk_enqueue_sweeps(int n) {
  const int k = blk_k, prec = blk_prec;
  const int64_t nl = P->nlam;
  const int* pending = &b_sc.p->pending;
  const int* active = &b_sc.p->active;
  for (int c = 0; c < n; c++) {
    // one PCPG iteration for every running column; the launches are skipped on
    // the device once no column runs (active == 0)
    bF(k, prec, b_w.p, b_Fw.p, nullptr, b_w.p, active);
    HC(launch_kb_reduce(st, b_nparts, k, b_part.p, b_sc.p, 0, active));
    HC(launch_kb_update(st, nl, k, 0, b_lam.p, b_g.p, b_w.p, b_Fw.p, b_sc.p, active));
    bPM(k, prec, b_g.p, b_z.p, b_hg.p, active);
    bP(k, b_z.p, b_Pz.p, b_h.p, b_g.p, active);
    HC(launch_kb_reduce(st, b_nparts, k, b_part.p, b_sc.p, 1, active));
    HC(launch_kb_update(st, nl, k, 1, b_w.p, nullptr, b_Pz.p, nullptr, b_sc.p, active));
    if (prec == 32) {
      // columns waiting at a refinement point: FP64 true residual g = F lam - d,
      // certification or restart from it (skipped on the device when none waits)
      bF(k, 64, b_lam.p, b_g64.p, b_d.p, nullptr, pending);
      bPM(k, prec, b_g64.p, b_z.p, b_hg.p, pending);
      bP(k, b_z.p, b_Pz.p, b_h.p, b_g64.p, pending);
      HC(launch_kb_reduce(st, b_nparts, k, b_part.p, b_sc.p, 3));
      HC(launch_kb_save_hg(st, 6 * P->nsub, k, b_sc.p, b_hg.p, b_hgf.p));
      HC(launch_kb_update(st, nl, k, 2, b_g.p, b_w.p, b_g64.p, b_Pz.p, b_sc.p, &b_sc.p->restart));
    }
  }
}

void Engine::blk_sync(BScalars* sh) {
  HC(hipMemcpyAsync(sh, b_sc.p, sizeof(BScalars), hipMemcpyDeviceToHost, st));
  HC(hipStreamSynchronize(st));
}

void Engine::blk_finish(const std::vector<int>& cols, double* u_h, double* lam_h, const BScalars* sh) {
  blk_recover(sh);
  blk_collect(cols, u_h, lam_h);
}

void Engine::blk_recover(const BScalars* sh) {
  const int N = P->nsub, k = blk_k, prec = blk_prec;
  const int64_t np = off_n[N], ng = 3 * (int64_t)P->nPoints;
  // alpha (H alpha per subdomain) from the final FP64 true residual's projection.
  // Columns certified on their FP64 true residual saved it at certification; it is
  // recomputed (the same arithmetic on the same frozen lambda) unless every stopped
  // column has it
  bool saved = sh != nullptr && prec == 32;
  for (int j = 0; saved && j < k; j++)
    if (sh->done[j] == 3 || sh->done[j] == 6 || (sh->done[j] == 2 && !sh->hgok[j])) saved = false;
  if (!saved) {
    bF(k, 64, b_lam.p, b_g64.p, b_d.p, nullptr);
    bPM(k, prec, b_g64.p, b_z.p, b_hg.p);
    bP(k, b_z.p, b_Pz.p, b_h.p, b_g64.p);
    HC(launch_kb_reduce(st, b_nparts, k, b_part.p, b_sc.p, 4));
  }
  const double* hal = saved ? b_hgf.p : b_hg.p;
  // recovery: u = K_c^+ (f - Bf^T lam) - Rf alpha (Cluster.cpp:1648-1678), on a
  // copy of the loads (the running columns keep theirs)
  HC(hipMemcpyAsync(b_V2.p, b_V.p, (size_t)np * k * sizeof(double), hipMemcpyDeviceToDevice, st));
  HC(launch_kb_sub_bt(st, N, maxns, k, np, d_nr.p, d_ns.p, d_perm1.p, d_off_n.p, d_off_b.p, d_gptr.p, d_gidx.p,
                      b_lam.p, b_V2.p));
  b_kplus(k, true, b_V2.p);
  HC(launch_kb_recover(st, N, maxn, k, KP, np, ng, d_n.p, d_off_n.p, d_R.p, d_off_R.p, b_arow.p, b_la.p, hal, b_X.p,
                       d_gdof.p, d_owner.p, b_U.p));
}

void Engine::blk_collect(const std::vector<int>& cols, double* u_h, double* lam_h) {
  const int k = blk_k;
  const int64_t nl = P->nlam, ng = 3 * (int64_t)P->nPoints;
  for (size_t c = 0; c < cols.size(); c++)
    if (u_h)
      HC(hipMemcpyAsync(u_h + c * (size_t)ng, b_U.p + (size_t)cols[c] * ng, ng * sizeof(double),
                        hipMemcpyDeviceToHost, st));
  if (lam_h) {
    std::vector<double> lt((size_t)nl * k);
    HC(hipMemcpyAsync(lt.data(), b_lam.p, lt.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HC(hipStreamSynchronize(st));
    for (size_t c = 0; c < cols.size(); c++)
      for (int64_t r = 0; r < nl; r++) lam_h[c * (size_t)nl + r] = lt[(size_t)r * k + cols[c]];
  }
  blk_set_status(cols, 4);
  HC(hipStreamSynchronize(st));
}

void Engine::solve_blocks(int nb, double eps, int max_iter, int k, int prec, const double* f_h, double* u_h,
                          int* iters, int* status, double* res, BlockInfo* info) {
  if (nb < 1) throw std::runtime_error("at least one block");
  blk_contexts(nb);
  const int64_t np = off_n[P->nsub], ng = 3 * (int64_t)P->nPoints;
  std::vector<int> cols(k);
  for (int j = 0; j < k; j++) cols[j] = j;
  for (int c = 0; c < nb; c++) {  // loads to HBM: outside the timed span
    use_ctx(c);
    blk_begin(k, eps, max_iter, prec);
    blk_load(cols, f_h + (size_t)c * k * np);
    HC(hipStreamSynchronize(st));
  }
  Events ev(nb + 4);
  // ev[0]: start (context 0), ev[1]: init done (context 0), ev[2]: loop done
  // (context 0), ev[3]: span end; ev[4 + c]: end of context c
  use_ctx(0);
  HC(hipEventRecord(ev[0], st));
  for (int c = 1; c < nb; c++) {
    use_ctx(c);
    HC(hipStreamWaitEvent(st, ev[0], 0));
  }
  for (int c = 0; c < nb; c++) {
    use_ctx(c);
    blk_init(cols);
    if (c == 0) HC(hipEventRecord(ev[1], st));
  }
  std::vector<BScalars> sh(nb);
  std::vector<int> sweeps(nb, 0), run(nb, 1);
  const int chunk = poll_chunk;
  for (bool any = true; any;) {
    any = false;
    for (int c = 0; c < nb; c++)  // one chunk per running block, all streams busy at once
      if (run[c]) {
        use_ctx(c);
        blk_sweeps(chunk);
        sweeps[c] += chunk;
      }
    for (int c = 0; c < nb; c++)
      if (run[c]) {
        use_ctx(c);
        blk_sync(&sh[c]);
        bool r = false;
        for (int j = 0; j < k; j++)
          if (sh[c].done[j] == 0 || sh[c].done[j] == 1) r = true;
        if (!r || sweeps[c] > max_iter + 2 * chunk) run[c] = 0;
        else any = true;
      }
  }
  use_ctx(0);
  HC(hipEventRecord(ev[2], st));
  for (int c = 0; c < nb; c++) {
    use_ctx(c);
    blk_recover(&sh[c]);
    if (c > 0) HC(hipEventRecord(ev[4 + c], st));
  }
  use_ctx(0);
  for (int c = 1; c < nb; c++) HC(hipStreamWaitEvent(st, ev[4 + c], 0));
  HC(hipEventRecord(ev[3], st));
  HC(hipEventSynchronize(ev[3]));
  for (int c = 0; c < nb; c++) {
    use_ctx(c);
    std::vector<int> its(k), stt(k);
    for (int j = 0; j < k; j++) {
      its[j] = sh[c].it[j];
      stt[j] = sh[c].done[j];
    }
    blk_collect(cols, u_h + (size_t)c * k * ng, nullptr);
    BScalars s2;
    blk_sync(&s2);
    for (int j = 0; j < k; j++) {
      const size_t o = (size_t)c * k + j;
      if (iters) iters[o] = its[j];
      int sj = stt[j];
      if (sj == 0 || sj == 1) sj = 3;
      if (status) status[o] = sj;
      if (res) res[o] = s2.res[j];
    }
  }
  use_ctx(0);
  if (info) {
    float ms = 0;
    HC(hipEventElapsedTime(&ms, ev[0], ev[3]));
    info->solve_ms = ms;
    HC(hipEventElapsedTime(&ms, ev[0], ev[1]));
    info->init_ms = ms;
    HC(hipEventElapsedTime(&ms, ev[1], ev[2]));
    info->pcpg_ms = ms;
    HC(hipEventElapsedTime(&ms, ev[2], ev[3]));
    info->recover_ms = ms;
    info->sweeps = *std::max_element(sweeps.begin(), sweeps.end());
    int cs = 0;
    for (int c = 0; c < nb; c++)
      for (int j = 0; j < k; j++) cs += sh[c].ncert[j];
    info->cert_steps = cs;
  }
}

void Engine::solve_block(double eps, int max_iter, int k, int prec, const double* f_h, double* u_h, double* lam_h,
                         int* iters, int* status, double* res, BlockInfo* info) {
  blk_begin(k, eps, max_iter, prec);
  std::vector<int> cols(k);
  for (int j = 0; j < k; j++) cols[j] = j;
  Events ev(4);
  blk_load(cols, f_h);  // host loads to HBM: outside the timed solve
  HC(hipEventRecord(ev[0], st));
  blk_init(cols);
  HC(hipEventRecord(ev[1], st));
  const int chunk = poll_chunk;
  int sweeps = 0;
  BScalars sh;
  for (;;) {
    blk_sweeps(chunk);
    sweeps += chunk;
    blk_sync(&sh);
    bool running = false;
    for (int j = 0; j < k; j++)
      if (sh.done[j] == 0 || sh.done[j] == 1) running = true;
    if (!running) break;
    if (sweeps > max_iter + 2 * chunk) break;
  }
  HC(hipEventRecord(ev[2], st));
  // per-column results before the slots are released
  std::vector<int> its(k), stt(k);
  for (int j = 0; j < k; j++) {
    its[j] = sh.it[j];
    stt[j] = sh.done[j];
  }
  blk_recover(&sh);
  HC(hipEventRecord(ev[3], st));
  blk_collect(cols, u_h, lam_h);  // displacements to the host: outside the timed solve
  blk_sync(&sh);
  for (int j = 0; j < k; j++) {
    if (iters) iters[j] = its[j];
    int s = stt[j];
    if (s == 0 || s == 1) s = 3;  // stopped by the sweep limit
    if (status) status[j] = s;
    if (res) res[j] = sh.res[j];
  }
  if (info) {
    float ms = 0;
    HC(hipEventElapsedTime(&ms, ev[0], ev[3]));
    info->solve_ms = ms;
    HC(hipEventElapsedTime(&ms, ev[0], ev[1]));
    info->init_ms = ms;
    HC(hipEventElapsedTime(&ms, ev[1], ev[2]));
    info->pcpg_ms = ms;
    HC(hipEventElapsedTime(&ms, ev[2], ev[3]));
    info->recover_ms = ms;
    info->sweeps = sweeps;
    int nc = 0;
    for (int j = 0; j < k; j++) nc += sh.ncert[j];
    info->cert_steps = nc;
  }
}

}  // namespace hfls
