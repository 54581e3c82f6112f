// gfx950 kernels of the HTFETI hot path.
//
// Setup (per subdomain, batched over a chunk of subdomains in one launch):
//   k_assemble   hex8 stiffness (Data.cpp:297-490) straight into a dense
//                interior-first matrix, lower triangle, 8-colour element order so
//                no two waves touch one entry (deterministic, no atomics)
//   k_panel_*    Cholesky of a <=64 wide diagonal block (+ its inverse) and TRSM of the rows below
//   k_syrk       trailing update of a right-looking partial Cholesky (matrix cores)
//   k_detect     kernel detection on the Schur complement of the s-set: cyclic
//                Jacobi, eigen-jump test (Matrix.cpp:2352-2368)
//   k_coldot / k_bwd / k_fwd   triangular solves with a few right-hand sides
//   k_gs_pivots  Gram-Schmidt (Matrix.cpp:2450-2472), fixing DOFs, regularisation
// Solve (PCPG, Cluster.cpp:1363-1458):
//   k_gemv_gather   y_i = A_i (B_i^T v): LDS-staged interface gather + coalesced
//                   column streaming of a dense per-subdomain operator (K+_bb or S_i)
//   k_corr_*, k_cluster_*   HTFETI cluster correction (Cluster.cpp:1129-1181)
//   k_scatter, k_gh, k_proj_* , k_reduce   dual-space vectors, projector, dots
// All reductions use fixed trees, so results are bitwise reproducible run to run.
#include "kernels.hpp"

namespace hfls {

#define DEV __device__ __forceinline__

static DEV double wave_sum(double v) {
  // fixed butterfly over 64 lanes
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NT>
static DEV double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NT / 64; i++) s += sh[i];
  __syncthreads();
  return s;
}

// ============================================================ assembly
__global__ __launch_bounds__(256) void k_assemble(const AsmSub* __restrict__ subs,
                                                  const int64_t* __restrict__ moff, const int* __restrict__ ld,
                                                  const int* __restrict__ pinv,
                                                  const int64_t* __restrict__ pinv_off, double* __restrict__ M,
                                                  double* __restrict__ f, const int64_t* __restrict__ f_off,
                                                  double* __restrict__ kdiag, double dx, double dy, double dz,
                                                  double nu, double ax, double ay, double az) {
  const int mi = blockIdx.x;
  const AsmSub s = subs[mi];
  double* A = M + moff[mi];
  const int64_t lda = ld[mi];
  const int* pv = pinv + pinv_off[mi];
  double* fd = f ? f + f_off[mi] : nullptr;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nnx = s.ne[0] + 1, nny = s.ne[1] + 1;
  const double PT = 0.577350269189626;
  const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oy[8] = {0, 0, 1, 1, 0, 0, 1, 1}, oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  const double acc[3] = {ax, ay, az};

  for (int color = 0; color < 8; color++) {
    const int cx = color & 1, cy = (color >> 1) & 1, cz = (color >> 2) & 1;
    const int ncx = (s.ne[0] - cx + 1) / 2, ncy = (s.ne[1] - cy + 1) / 2, ncz = (s.ne[2] - cz + 1) / 2;
    const int cnt = ncx * ncy * ncz;
    for (int q = wave; q < cnt; q += 4) {
      const int ex = cx + 2 * (q % ncx), ey = cy + 2 * ((q / ncx) % ncy), ez = cz + 2 * (q / (ncx * ncy));
      double Ee = s.E;
      if (s.mn[0] > 0) {
        const int pid = (s.e0[0] + ex) / s.mn[0] + ((s.e0[1] + ey) / s.mn[1]) * s.mN[0] +
                        ((s.e0[2] + ez) / s.mn[2]) * s.mN[0] * s.mN[1];
        if (pid & 1) Ee = s.E1;
      }
      const double c0 = Ee / ((1.0 + nu) * (1.0 - 2.0 * nu));
      const double l2 = (1.0 - nu) * c0, l1 = nu * c0, G = (0.5 - nu) * c0;
      double X[8], Y[8], Z[8];
      int ln[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int li = ex + ox[k], lj = ey + oy[k], lk = ez + oz[k];
        ln[k] = li + lj * nnx + lk * nnx * nny;
        // Mesh.cpp:158-168: x = (ii*dx - 0.5) + 0.5, z = kk*dz + 0
        X[k] = __dadd_rn(__dsub_rn(__dmul_rn((double)(s.e0[0] + li), dx), 0.5), 0.5);
        Y[k] = __dadd_rn(__dsub_rn(__dmul_rn((double)(s.e0[1] + lj), dy), 0.5), 0.5);
        Z[k] = __dmul_rn((double)(s.e0[2] + lk), dz);
      }
      double kij[9];
      int I9[9], J9[9];
#pragma unroll
      for (int t = 0; t < 9; t++) {
        kij[t] = 0.0;
        const int idx = lane + 64 * t;
        I9[t] = idx % 24;
        J9[t] = idx / 24;
      }
      double fe = 0.0;
      for (int gp = 0; gp < 8; gp++) {
        // gauss point order of Data.cpp:372-396
        const double r = (gp < 4) ? -PT : PT;
        const double sg = ((gp >> 1) & 1) ? PT : -PT;
        const double tg = (gp & 1) ? PT : -PT;
        double dNr[8], dNs[8], dNt[8], N[8];
        dNr[0] = 0.125 * -(1. - sg) * (1. - tg); dNr[1] = 0.125 * (1. - sg) * (1. - tg);
        dNr[2] = 0.125 * (1. + sg) * (1. - tg);  dNr[3] = 0.125 * -(1. + sg) * (1. - tg);
        dNr[4] = 0.125 * -(1. - sg) * (1. + tg); dNr[5] = 0.125 * (1. - sg) * (1. + tg);
        dNr[6] = 0.125 * (1. + sg) * (1. + tg);  dNr[7] = 0.125 * -(1. + sg) * (1. + tg);
        dNs[0] = 0.125 * -(1. - r) * (1. - tg);  dNs[1] = 0.125 * -(1. + r) * (1. - tg);
        dNs[2] = 0.125 * (1. + r) * (1. - tg);   dNs[3] = 0.125 * (1. - r) * (1. - tg);
        dNs[4] = 0.125 * -(1. - r) * (1. + tg);  dNs[5] = 0.125 * -(1. + r) * (1. + tg);
        dNs[6] = 0.125 * (1. + r) * (1. + tg);   dNs[7] = 0.125 * (1. - r) * (1. + tg);
        dNt[0] = 0.125 * -(1. - r) * (1. - sg);  dNt[1] = 0.125 * -(1. + r) * (1. - sg);
        dNt[2] = 0.125 * -(1. + r) * (1. + sg);  dNt[3] = 0.125 * -(1. - r) * (1. + sg);
        dNt[4] = 0.125 * (1. - r) * (1. - sg);   dNt[5] = 0.125 * (1. + r) * (1. - sg);
        dNt[6] = 0.125 * (1. + r) * (1. + sg);   dNt[7] = 0.125 * (1. - r) * (1. + sg);
        N[0] = 0.125 * (1. - r) * (1. - sg) * (1. - tg); N[1] = 0.125 * (r + 1.) * (1. - sg) * (1. - tg);
        N[2] = 0.125 * (r + 1.) * (sg + 1.) * (1. - tg); N[3] = 0.125 * (1. - r) * (sg + 1.) * (1. - tg);
        N[4] = 0.125 * (1. - r) * (1. - sg) * (tg + 1.); N[5] = 0.125 * (r + 1.) * (1. - sg) * (tg + 1.);
        N[6] = 0.125 * (r + 1.) * (sg + 1.) * (tg + 1.); N[7] = 0.125 * (1. - r) * (sg + 1.) * (tg + 1.);
        double Jm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 8; k++) {
          Jm[0] += dNr[k] * X[k]; Jm[1] += dNs[k] * X[k]; Jm[2] += dNt[k] * X[k];
          Jm[3] += dNr[k] * Y[k]; Jm[4] += dNs[k] * Y[k]; Jm[5] += dNt[k] * Y[k];
          Jm[6] += dNr[k] * Z[k]; Jm[7] += dNs[k] * Z[k]; Jm[8] += dNt[k] * Z[k];
        }
        // Data.cpp:276-294
        const double det = Jm[0] * (Jm[4] * Jm[8] - Jm[5] * Jm[7]) - Jm[3] * (Jm[1] * Jm[8] - Jm[7] * Jm[2]) +
                           Jm[6] * (Jm[1] * Jm[5] - Jm[4] * Jm[2]);
        const double id = 1.0 / det;
        double iJ[9];
        iJ[0] = (Jm[4] * Jm[8] - Jm[5] * Jm[7]) * id;
        iJ[1] = -(Jm[3] * Jm[8] - Jm[6] * Jm[5]) * id;
        iJ[2] = (Jm[3] * Jm[7] - Jm[6] * Jm[4]) * id;
        iJ[3] = -(Jm[1] * Jm[8] - Jm[7] * Jm[2]) * id;
        iJ[4] = (Jm[0] * Jm[8] - Jm[6] * Jm[2]) * id;
        iJ[5] = -(Jm[0] * Jm[7] - Jm[1] * Jm[6]) * id;
        iJ[6] = (Jm[1] * Jm[5] - Jm[2] * Jm[4]) * id;
        iJ[7] = -(Jm[0] * Jm[5] - Jm[2] * Jm[3]) * id;
        iJ[8] = (Jm[0] * Jm[4] - Jm[1] * Jm[3]) * id;
        double dNx[8], dNy[8], dNz[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          dNx[k] = iJ[0] * dNr[k] + iJ[3] * dNs[k] + iJ[6] * dNt[k];
          dNy[k] = iJ[1] * dNr[k] + iJ[4] * dNs[k] + iJ[7] * dNt[k];
          dNz[k] = iJ[2] * dNr[k] + iJ[5] * dNs[k] + iJ[8] * dNt[k];
        }
#pragma unroll
        for (int t = 0; t < 9; t++) {
          const int I = I9[t], J = J9[t];
          const int kI = I & 7, cI = I >> 3, kJ = J & 7, cJ = J >> 3;
          // columns of the strain-displacement matrix Gt (Data.cpp:451-464)
          double gi[6], gj[6];
          gi[0] = cI == 0 ? dNx[kI] : 0.0; gi[1] = cI == 1 ? dNy[kI] : 0.0; gi[2] = cI == 2 ? dNz[kI] : 0.0;
          gi[3] = cI == 0 ? dNy[kI] : (cI == 1 ? dNx[kI] : 0.0);
          gi[4] = cI == 1 ? dNz[kI] : (cI == 2 ? dNy[kI] : 0.0);
          gi[5] = cI == 0 ? dNz[kI] : (cI == 2 ? dNx[kI] : 0.0);
          gj[0] = cJ == 0 ? dNx[kJ] : 0.0; gj[1] = cJ == 1 ? dNy[kJ] : 0.0; gj[2] = cJ == 2 ? dNz[kJ] : 0.0;
          gj[3] = cJ == 0 ? dNy[kJ] : (cJ == 1 ? dNx[kJ] : 0.0);
          gj[4] = cJ == 1 ? dNz[kJ] : (cJ == 2 ? dNy[kJ] : 0.0);
          gj[5] = cJ == 0 ? dNz[kJ] : (cJ == 2 ? dNx[kJ] : 0.0);
          const double cg0 = l2 * gj[0] + l1 * gj[1] + l1 * gj[2];
          const double cg1 = l1 * gj[0] + l2 * gj[1] + l1 * gj[2];
          const double cg2 = l1 * gj[0] + l1 * gj[1] + l2 * gj[2];
          const double v = gi[0] * cg0 + gi[1] * cg1 + gi[2] * cg2 + G * (gi[3] * gj[3] + gi[4] * gj[4] + gi[5] * gj[5]);
          kij[t] += v * det;
        }
        if (lane < 24) fe += acc[lane >> 3] * N[lane & 7] * det;
      }
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int I = I9[t], J = J9[t];
        const int dI = 3 * ln[I & 7] + (I >> 3), dJ = 3 * ln[J & 7] + (J >> 3);
        const int pI = pv[dI], pJ = pv[dJ];
        if (pI >= pJ) A[pI + (int64_t)pJ * lda] += kij[t];
      }
      if (fd && lane < 24) fd[3 * ln[lane & 7] + (lane >> 3)] += fe;
    }
    __syncthreads();
  }
  if (kdiag) {
    double* kd = kdiag + f_off[mi];
    for (int l = threadIdx.x; l < s.n; l += 256) {
      const int p = pv[l];
      kd[l] = A[p + (int64_t)p * lda];
    }
  }
  // augmented interior block [K_rr K_rb I; K_br K_bb 0; I 0 0]
  for (int i = threadIdx.x; i < s.nr; i += 256) A[(s.n + i) + (int64_t)i * lda] = 1.0;
}

hipError_t launch_assemble(hipStream_t st, int nmat, const AsmSub* subs, const int64_t* moff, const int* ld,
                           const int* pinv, const int64_t* pinv_off, double* M, double* f, const int64_t* f_off,
                           double* kdiag, double dx, double dy, double dz, double nu, double ax, double ay,
                           double az) {
  hipLaunchKernelGGL(k_assemble, dim3(nmat), dim3(256), 0, st, subs, moff, ld, pinv, pinv_off, M, f, f_off, kdiag,
                     dx, dy, dz, nu, ax, ay, az);
  return hipGetLastError();
}

// ============================================================ Cholesky panel
// k_panel_diag: one workgroup (64 threads) per matrix factors the <=64 wide
// diagonal block in LDS, writes it back, and writes the inverse of the factor,
// W = L^-1 (lower, column major, zero outside w x w), for the TRSM below.
// L[c*64 + r] holds the diagonal block (column major).  The factorisation is left
// looking: thread t forms L[t][j] = (A[t][j] - sum_{k<j} L[t][k] L[j][k]) / L[j][j]
// from LDS reads only (no store inside the sum, so its loads pipeline); the
// products are subtracted from A[t][j] in ascending k, the order of a
// right-looking update.
__global__ __launch_bounds__(64) void k_panel_diag(const PanelDesc* __restrict__ desc,
                                                   const int64_t* __restrict__ moff, const int* __restrict__ ld,
                                                   double* __restrict__ M, double* __restrict__ Winv,
                                                   int* __restrict__ err) {
  const int mi = blockIdx.y;
  const PanelDesc pd = desc[mi];
  double* W = Winv + (int64_t)mi * 4096;
  const int t = threadIdx.x;
  if (!pd.active || pd.w <= 0) return;
  const int k0 = pd.k0, w = pd.w;
  double* A = M + moff[mi];
  const int64_t lda = ld[mi];
  __shared__ double L[64 * 64];
  __shared__ double Y[64 * 65];
  {  // all 64 column loads in flight before the LDS stores
    double col[64];
#pragma unroll
    for (int c = 0; c < 64; c++) col[c] = (c < w && t < w && t >= c) ? A[(k0 + t) + (int64_t)(k0 + c) * lda] : 0.0;
#pragma unroll
    for (int c = 0; c < 64; c++)
      if (c < w) L[c * 64 + t] = col[c];
  }
  __syncthreads();
  for (int j = 0; j < w; j++) {
    double sj = 0.0;
    if (t >= j && t < w) {
      sj = L[j * 64 + t];
      for (int k = 0; k < j; k++) sj -= L[k * 64 + t] * L[k * 64 + j];
    }
    if (t == j) {
      double dj = 1.0;
      if (sj > 0.0) dj = sqrt(sj);
      else atomicOr(err, 1);
      L[j * 64 + j] = dj;
    }
    __syncthreads();
    if (t > j && t < w) L[j * 64 + t] = sj / L[j * 64 + j];
    __syncthreads();
  }
  if (t < w)
    for (int c = 0; c <= t; c++) A[(k0 + t) + (int64_t)(k0 + c) * lda] = L[c * 64 + t];
  // column t of L^-1 by forward substitution (thread t), kept in Y[t*65 + i]
  for (int i = 0; i < 64; i++) {
    double y = 0.0;
    if (t < w && i < w && i >= t) {
      double s = i == t ? 1.0 : 0.0;
      for (int k = t; k < i; k++) s -= L[k * 64 + i] * Y[t * 65 + k];
      y = s / L[i * 64 + i];
    }
    Y[t * 65 + i] = y;
  }
  __syncthreads();
  for (int c = 0; c < 64; c++) W[c * 64 + t] = Y[c * 65 + t];
}

// k_panel_trsm: the rows below the diagonal block, X L^T = A_rows, as X = A_rows W^T
// on the matrix cores: one workgroup (256 threads) per 64 rows, the product laid out
// as in k_syrk (D columns are matrix rows, so the stores are row-contiguous).
typedef double trsm_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_panel_trsm(const PanelDesc* __restrict__ desc,
                                                    const int64_t* __restrict__ moff, const int* __restrict__ ld,
                                                    double* __restrict__ M, const double* __restrict__ Winv) {
  const int mi = blockIdx.y;
  const PanelDesc pd = desc[mi];
  if (!pd.active || pd.w <= 0) return;
  const int k0 = pd.k0, w = pd.w, end = pd.end;
  const int ri = k0 + w + blockIdx.x * 64;
  if (ri >= end) return;
  double* A = M + moff[mi];
  const int64_t lda = ld[mi];
  const double* W = Winv + (int64_t)mi * 4096;
  __shared__ double As[64 * 65];  // [k][row] of the rows being solved
  __shared__ double Bs[64 * 65];  // [k][c] = W[c][k]
#pragma unroll
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int k = idx >> 6, r = idx & 63;
    As[k * 65 + r] = (k < w && ri + r < end) ? A[(ri + r) + (int64_t)(k0 + k) * lda] : 0.0;
    Bs[k * 65 + r] = W[k * 64 + r];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  trsm_d4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[c][r] = 0.0;
  const int kw = (w + 3) & ~3;
  for (int kk = 0; kk < kw; kk += 4) {
    const double a = Bs[(kk + (lane >> 4)) * 65 + 16 * v + (lane & 15)];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const double b = As[(kk + (lane >> 4)) * 65 + 16 * c + (lane & 15)];
      acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int row = ri + 16 * c + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = 16 * v + 4 * r + (lane >> 4);
      if (row < end && col < w) A[row + (int64_t)(k0 + col) * lda] = acc[c][r];
    }
  }
}

hipError_t launch_panel(hipStream_t st, int nmat, int maxrows, const PanelDesc* desc, const int64_t* moff,
                        const int* ld, double* M, double* Winv, int* err) {
  hipLaunchKernelGGL(k_panel_diag, dim3(1, nmat), dim3(64), 0, st, desc, moff, ld, M, Winv, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || maxrows <= 0) return e;
  hipLaunchKernelGGL(k_panel_trsm, dim3((maxrows + 63) / 64, nmat), dim3(256), 0, st, desc, moff, ld, M, Winv);
  return hipGetLastError();
}

// ============================================================ SYRK trailing update
// C[r][c] -= sum_{k in [k0, k0+w)} L[r][k] L[c][k] for t0 <= c <= r, r < end, c < cend:
// the lower trapezoid of a right-looking partial Cholesky step (cend = end: the
// whole trailing triangle; cend < end: only the columns of the current super-panel).
// One workgroup per 64x64 tile (ti, tj), tj < ceil((cend - t0)/64), ti >= tj,
// enumerated column by column.  K is streamed in chunks of 32 through LDS
// ([k][row], padded rows); the product runs on the matrix cores
// (v_mfma_f64_16x16x4f64, lane layouts A[l%16][l/16], B[l/16][l%16], D row 4r+l/16,
// col l%16) as C^T blocks, so that a D column is a matrix row and the
// read-modify-write of C is 16 consecutive doubles per lane group.  Wave v owns
// the matrix columns 16v..16v+15 of the tile and all 64 rows.
typedef double syrk_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_syrk(const SyrkDesc* __restrict__ desc, const int64_t* __restrict__ moff,
                                              const int* __restrict__ ld, double* __restrict__ M) {
  const int mi = blockIdx.y;
  const SyrkDesc sd = desc[mi];
  if (!sd.active || sd.w <= 0) return;
  const int k0 = sd.k0, w = sd.w, t0 = sd.t0, cend = sd.cend, end = sd.end;
  const int m = end - t0;
  if (m <= 0 || cend <= t0) return;
  const int T = (m + 63) / 64, Tc = min(T, (cend - t0 + 63) / 64);
  const int64_t q = blockIdx.x;
  auto count = [&](int64_t x) { return x * T - x * (x - 1) / 2; };  // tiles in columns < x
  if (q >= count(Tc)) return;
  const double tb = T + 0.5;
  int tj = (int)(tb - sqrt(tb * tb - 2.0 * (double)q));
  tj = max(0, min(tj, Tc - 1));
  while (tj > 0 && count(tj) > q) tj--;
  while (tj + 1 < Tc && count(tj + 1) <= q) tj++;
  const int ti = tj + (int)(q - count(tj));
  const int ri = t0 + 64 * ti, cj = t0 + 64 * tj;
  double* A = M + moff[mi];
  const int64_t lda = ld[mi];
  __shared__ double As[32 * 65];  // rows of the tile's row block
  __shared__ double Bs[32 * 65];  // rows of the tile's column block
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  syrk_d4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[c][r] = 0.0;
  // the next K chunk is loaded into registers while the matrix cores work on the
  // current one from LDS
  double pa[8], pb[8];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int idx = threadIdx.x + 256 * t, k = idx >> 6, r = idx & 63;
      const bool kin = kc + k < w;
      const int64_t col = k0 + kc + k;
      pa[t] = (kin && ri + r < end) ? A[(ri + r) + col * lda] : 0.0;
      pb[t] = (kin && cj + r < end) ? A[(cj + r) + col * lda] : 0.0;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int idx = threadIdx.x + 256 * t, k = idx >> 6, r = idx & 63;
      As[k * 65 + r] = pa[t];
      Bs[k * 65 + r] = pb[t];
    }
  };
  fetch(0);
  stash();
  __syncthreads();
  for (int kc = 0; kc < w; kc += 32) {
    const bool more = kc + 32 < w;
    if (more) fetch(kc + 32);
    const int kw = min(32, (w - kc + 3) & ~3);
    for (int kk = 0; kk < kw; kk += 4) {
      const double a = Bs[(kk + (lane >> 4)) * 65 + 16 * v + (lane & 15)];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double b = As[(kk + (lane >> 4)) * 65 + 16 * c + (lane & 15)];
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
      }
    }
    __syncthreads();
    if (more) {
      stash();
      __syncthreads();
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int row = ri + 16 * c + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = cj + 16 * v + 4 * r + (lane >> 4);
      if (row < end && col < cend && row >= col) A[row + (int64_t)col * lda] -= acc[c][r];
    }
  }
}

hipError_t launch_syrk(hipStream_t st, int nmat, int64_t maxtiles, const SyrkDesc* desc, const int64_t* moff,
                       const int* ld, double* M) {
  if (maxtiles < 1) return hipSuccess;
  if (maxtiles > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_syrk, dim3((unsigned)maxtiles, nmat), dim3(256), 0, st, desc, moff, ld, M);
  return hipGetLastError();
}

// ============================================================ block copy
// dst[p + q*dld] = scale * src(srow0+max(p,q), srow0+min(p,q)) for p,q < n
// (symmetric fill from a lower triangle); aug != 0 also writes the identity
// coupling dst[(n+p) + p*dld] = 1 of the augmented matrix [S I; I 0].
__global__ __launch_bounds__(256) void k_copy_block(const double* __restrict__ src, const int64_t* __restrict__ soff,
                                                    const int* __restrict__ sld, const int* __restrict__ srow0,
                                                    const int* __restrict__ nn, double* __restrict__ dst,
                                                    const int64_t* __restrict__ doff, const int* __restrict__ dld,
                                                    double scale, int aug) {
  const int mi = blockIdx.y;
  const int n = nn[mi];
  const int q = blockIdx.x;
  if (q >= n) return;
  const double* S = src + soff[mi];
  const int64_t ls = sld[mi];
  const int r0 = srow0[mi];
  double* D = dst + doff[mi];
  const int64_t ldd = dld[mi];
  for (int p = threadIdx.x; p < n; p += 256) {
    const int hi = p > q ? p : q, lo = p > q ? q : p;
    D[p + (int64_t)q * ldd] = scale * S[(r0 + hi) + (int64_t)(r0 + lo) * ls];
  }
  if (aug && threadIdx.x == 0) D[(n + q) + (int64_t)q * ldd] = 1.0;
}

hipError_t launch_copy_block(hipStream_t st, int nmat, int maxn, const double* src, const int64_t* soff,
                             const int* sld, const int* srow0, const int* n, double* dst, const int64_t* doff,
                             const int* dld, double scale, int aug) {
  hipLaunchKernelGGL(k_copy_block, dim3(maxn, nmat), dim3(256), 0, st, src, soff, sld, srow0, n, dst, doff, dld,
                     scale, aug);
  return hipGetLastError();
}

// ============================================================ kernel detection
// Cyclic Jacobi on the diagonally scaled Schur complement of the s-set
// (Matrix.cpp:2137-2149 scaling, 2352 eigen-decomposition, 2359-2368 jump test,
// 2393-2421 R_s).  One workgroup per subdomain, 7 <= m <= 32.
__global__ __launch_bounds__(64) void k_detect(const double* __restrict__ M, const int64_t* __restrict__ moff,
                                               const int* __restrict__ ld, const int* __restrict__ row0,
                                               const int* __restrict__ mm, const double* __restrict__ kdiag,
                                               const int64_t* __restrict__ kd_off, const int* __restrict__ perm,
                                               const int* __restrict__ prow0, double* __restrict__ Rs,
                                               const int64_t* __restrict__ rs_off, const int* __restrict__ ldr,
                                               int* __restrict__ defect) {
  const int mi = blockIdx.x;
  const int m = mm[mi];
  const double* A = M + moff[mi];
  const int64_t lda = ld[mi];
  const int r0 = row0[mi];
  __shared__ double S[32 * 65];
  __shared__ double V[32 * 65];
  __shared__ double cs[3];
  __shared__ double W[32];
  __shared__ double D[32];
  __shared__ int idx[32];
  const int t = threadIdx.x;
  if (t < m) D[t] = kdiag[kd_off[mi] + perm[kd_off[mi] + prow0[mi] + t]];
  __syncthreads();
  if (t < m) {
    for (int j = 0; j < m; j++) {
      const int hi = t > j ? t : j, lo = t > j ? j : t;
      S[t * 65 + j] = A[(r0 + hi) + (int64_t)(r0 + lo) * lda] / sqrt(D[t] * D[j]);
      V[t * 65 + j] = (t == j) ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  for (int sweep = 0; sweep < 60; sweep++) {
    if (t == 0) {
      double off = 0.0, tot = 0.0;
      for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) {
          const double v = S[i * 65 + j] * S[i * 65 + j];
          tot += v;
          if (i != j) off += v;
        }
      cs[2] = (off <= 1e-30 * tot) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (cs[2] != 0.0) break;
    for (int p = 0; p < m - 1; p++)
      for (int q = p + 1; q < m; q++) {
        if (t == 0) {
          const double apq = S[p * 65 + q];
          double c = 1.0, sn = 0.0;
          if (apq != 0.0) {
            const double tau = (S[q * 65 + q] - S[p * 65 + p]) / (2.0 * apq);
            const double tt = (tau >= 0.0) ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tt * tt);
            sn = tt * c;
          }
          cs[0] = c;
          cs[1] = sn;
        }
        __syncthreads();
        const double c = cs[0], sn = cs[1];
        if (sn != 0.0) {
          if (t < m) {
            const double akp = S[t * 65 + p], akq = S[t * 65 + q];
            S[t * 65 + p] = c * akp - sn * akq;
            S[t * 65 + q] = sn * akp + c * akq;
          }
          __syncthreads();
          if (t < m) {
            const double apk = S[p * 65 + t], aqk = S[q * 65 + t];
            S[p * 65 + t] = c * apk - sn * aqk;
            S[q * 65 + t] = sn * apk + c * aqk;
            const double vkp = V[t * 65 + p], vkq = V[t * 65 + q];
            V[t * 65 + p] = c * vkp - sn * vkq;
            V[t * 65 + q] = sn * vkp + c * vkq;
          }
        }
        __syncthreads();
      }
  }
  if (t == 0) {
    for (int i = 0; i < m; i++) {
      W[i] = S[i * 65 + i];
      idx[i] = i;
    }
    for (int i = 0; i < m; i++)  // selection sort ascending
      for (int j = i + 1; j < m; j++)
        if (W[j] < W[i]) {
          const double tw = W[i]; W[i] = W[j]; W[j] = tw;
          const int ti = idx[i]; idx[i] = idx[j]; idx[j] = ti;
        }
    const int itmax = m < 20 ? m : 20;
    int dfc = 0;
    for (int i = itmax - 1; i > 0; i--)
      if (fabs(W[i - 1] / W[i]) < 1.0e-5) {
        dfc = i;
        break;
      }
    defect[mi] = dfc;
  }
  __syncthreads();
  double* R = Rs + rs_off[mi];
  const int64_t lr = ldr[mi];
  if (t < m)
    for (int k = 0; k < 6; k++) R[t + k * lr] = V[t * 65 + idx[k]] / sqrt(D[t]);
}

hipError_t launch_detect(hipStream_t st, int nmat, const double* M, const int64_t* moff, const int* ld,
                         const int* row0, const int* m, const double* kdiag, const int64_t* kd_off,
                         const int* perm, const int* prow0, double* Rs, const int64_t* rs_off, const int* ldr,
                         int* defect) {
  hipLaunchKernelGGL(k_detect, dim3(nmat), dim3(64), 0, st, M, moff, ld, row0, m, kdiag, kd_off, perm, prow0, Rs,
                     rs_off, ldr, defect);
  return hipGetLastError();
}

// ============================================================ column dots
// out[i + o*ldo] = alpha * sum_q L[(row0+q) + i*ld] * Y[q + o*ldy], i < ncols, o < nrhs <= 8
__global__ __launch_bounds__(256) void k_coldot(const double* __restrict__ L, const int64_t* __restrict__ loff,
                                                const int* __restrict__ ld, const int* __restrict__ row0,
                                                const int* __restrict__ mrows, const int* __restrict__ ncols,
                                                const double* __restrict__ Y, const int64_t* __restrict__ yoff,
                                                const int* __restrict__ ldy, double* __restrict__ out,
                                                const int64_t* __restrict__ ooff, const int* __restrict__ ldo,
                                                int nrhs, double alpha, double beta) {
  const int mi = blockIdx.y;
  const int nc = ncols[mi], mr = mrows[mi];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double* Lm = L + loff[mi];
  const int64_t l = ld[mi];
  const int r0 = row0[mi];
  const double* Ym = Y + yoff[mi];
  const int64_t ly = ldy[mi];
  double* O = out + ooff[mi];
  const int64_t lo = ldo[mi];
  for (int i = blockIdx.x * 16 + wv; i < nc && i < (int)(blockIdx.x + 1) * 16; i += 4) {
    double acc[8];
#pragma unroll
    for (int o = 0; o < 8; o++) acc[o] = 0.0;
    for (int q = lane; q < mr; q += 64) {
      const double a = Lm[(r0 + q) + (int64_t)i * l];
#pragma unroll
      for (int o = 0; o < 8; o++)
        if (o < nrhs) acc[o] += a * Ym[q + o * ly];
    }
#pragma unroll
    for (int o = 0; o < 8; o++)
      if (o < nrhs) {
        const double s = wave_sum(acc[o]);
        if (lane == 0) O[i + o * lo] = alpha * s + (beta != 0.0 ? beta * O[i + o * lo] : 0.0);
      }
  }
}

hipError_t launch_coldot(hipStream_t st, int nmat, int maxcols, const double* L, const int64_t* loff,
                         const int* ld, const int* row0, const int* mrows, const int* ncols, const double* Y,
                         const int64_t* yoff, const int* ldy, double* out, const int64_t* ooff, const int* ldo,
                         int nrhs, double alpha, double beta) {
  if (nrhs > 8) return hipErrorInvalidValue;
  int gx = (maxcols + 15) / 16;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(k_coldot, dim3(gx, nmat), dim3(256), 0, st, L, loff, ld, row0, mrows, ncols, Y, yoff, ldy, out,
                     ooff, ldo, nrhs, alpha, beta);
  return hipGetLastError();
}

// ============================================================ triangular solves
// Backward: solve L^T X = B in place (L lower, m x m, column major), nrhs <= 8.
__global__ __launch_bounds__(1024) void k_bwd(const double* __restrict__ L, const int64_t* __restrict__ loff,
                                             const int* __restrict__ ld, const int* __restrict__ mm,
                                             double* __restrict__ X, const int64_t* __restrict__ xoff,
                                             const int* __restrict__ ldx, int nrhs) {
  const int mi = blockIdx.x;
  const int m = mm[mi];
  const double* Lm = L + loff[mi];
  const int64_t l = ld[mi];
  double* Xm = X + xoff[mi];
  const int64_t lx = ldx[mi];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ double Lb[64 * 64];
  for (int j1 = m; j1 > 0; j1 -= 64) {
    const int j0 = j1 > 64 ? j1 - 64 : 0;
    const int bw = j1 - j0;
    // X[i] -= sum_{r >= j1} L[r,i] X[r]   (column i of L below the block)
    for (int i = j0 + wv; i < j1; i += 16) {
      double acc[8];
#pragma unroll
      for (int o = 0; o < 8; o++) acc[o] = 0.0;
      // four rows per lane and step, their loads issued together; a lane adds its
      // rows r, r + 64, r + 128, ... in ascending order
      for (int r = j1 + lane; r < m; r += 256) {
        double a[4];
#pragma unroll
        for (int u = 0; u < 4; u++) a[u] = r + 64 * u < m ? Lm[r + 64 * u + (int64_t)i * l] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (r + 64 * u < m)
#pragma unroll
            for (int o = 0; o < 8; o++)
              if (o < nrhs) acc[o] += a[u] * Xm[r + 64 * u + o * lx];
      }
#pragma unroll
      for (int o = 0; o < 8; o++)
        if (o < nrhs) {
          const double s = wave_sum(acc[o]);
          if (lane == 0) Xm[i + o * lx] -= s;
        }
    }
    for (int idx = threadIdx.x; idx < 64 * 64; idx += 1024) {
      const int c = idx >> 6, r = idx & 63;
      Lb[idx] = (c < bw && r < bw && r >= c) ? Lm[(j0 + r) + (int64_t)(j0 + c) * l] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < nrhs) {
      const int o = threadIdx.x;
      for (int i = bw - 1; i >= 0; i--) {
        double v = Xm[(j0 + i) + o * lx];
        for (int r = i + 1; r < bw; r++) v -= Lb[i * 64 + r] * Xm[(j0 + r) + o * lx];
        Xm[(j0 + i) + o * lx] = v / Lb[i * 64 + i];
      }
    }
    __syncthreads();
  }
}

hipError_t launch_bwd(hipStream_t st, int nmat, const double* L, const int64_t* loff, const int* ld,
                      const int* m, double* X, const int64_t* xoff, const int* ldx, int nrhs) {
  if (nrhs > 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bwd, dim3(nmat), dim3(1024), 0, st, L, loff, ld, m, X, xoff, ldx, nrhs);
  return hipGetLastError();
}

// Forward: solve L Y = B for the first m rows and update rows m..mall-1 with
// B_rest -= L_rest,1:m Y (right-looking), in place, nrhs <= 8.
__global__ __launch_bounds__(256) void k_fwd(const double* __restrict__ L, const int64_t* __restrict__ loff,
                                             const int* __restrict__ ld, const int* __restrict__ mm,
                                             const int* __restrict__ mall, double* __restrict__ X,
                                             const int64_t* __restrict__ xoff, const int* __restrict__ ldx,
                                             int nrhs) {
  const int mi = blockIdx.x;
  const int m = mm[mi], ma = mall[mi];
  const double* Lm = L + loff[mi];
  const int64_t l = ld[mi];
  double* Xm = X + xoff[mi];
  const int64_t lx = ldx[mi];
  __shared__ double Lb[64 * 64];
  __shared__ double Yb[64 * 8];
  for (int j0 = 0; j0 < m; j0 += 64) {
    const int bw = (m - j0) < 64 ? (m - j0) : 64;
    for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
      const int c = idx >> 6, r = idx & 63;
      Lb[idx] = (c < bw && r < bw && r >= c) ? Lm[(j0 + r) + (int64_t)(j0 + c) * l] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < nrhs) {
      const int o = threadIdx.x;
      for (int i = 0; i < bw; i++) {
        double v = Xm[(j0 + i) + o * lx];
        for (int c = 0; c < i; c++) v -= Lb[c * 64 + i] * Yb[c * 8 + o];
        v /= Lb[i * 64 + i];
        Yb[i * 8 + o] = v;
        Xm[(j0 + i) + o * lx] = v;
      }
    }
    __syncthreads();
    for (int r = j0 + bw + threadIdx.x; r < ma; r += 256) {
      double acc[8];
#pragma unroll
      for (int o = 0; o < 8; o++) acc[o] = 0.0;
      for (int c = 0; c < bw; c++) {
        const double a = Lm[r + (int64_t)(j0 + c) * l];
#pragma unroll
        for (int o = 0; o < 8; o++)
          if (o < nrhs) acc[o] += a * Yb[c * 8 + o];
      }
#pragma unroll
      for (int o = 0; o < 8; o++)
        if (o < nrhs) Xm[r + o * lx] -= acc[o];
    }
    __syncthreads();
  }
}

hipError_t launch_fwd(hipStream_t st, int nmat, const double* L, const int64_t* loff, const int* ld,
                      const int* m, const int* mall, double* X, const int64_t* xoff, const int* ldx, int nrhs) {
  if (nrhs > 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fwd, dim3(nmat), dim3(256), 0, st, L, loff, ld, m, mall, X, xoff, ldx, nrhs);
  return hipGetLastError();
}

// ============================================================ Gram-Schmidt + fixing DOFs
// R1 (n x 6, rows in interior-first order) -> modified Gram-Schmidt
// (Matrix.cpp:2450-2472); scatter to local order R and to the b-order block Rb;
// fixing DOFs = pivoted Cholesky of Rs Rs^T on the s-set (ties within 1e-9 go to
// the lowest index); regularise M2 at those diagonal entries with K_pp
// (Matrix.cpp:1033-1046).
__global__ __launch_bounds__(256) void k_gs_pivots(double* __restrict__ R, const int64_t* __restrict__ roff,
                                                   const int* __restrict__ nn, const int* __restrict__ perm,
                                                   const int64_t* __restrict__ perm_off, const int* __restrict__ nr,
                                                   const int* __restrict__ ns, const int* __restrict__ nbs,
                                                   double* __restrict__ Rb, const int64_t* __restrict__ rb_off,
                                                   double* __restrict__ R1, const int64_t* __restrict__ r1_off,
                                                   const double* __restrict__ kdiag,
                                                   const int64_t* __restrict__ kd_off, double* __restrict__ M2,
                                                   const int64_t* __restrict__ m2off, const int* __restrict__ ld2,
                                                   int* __restrict__ fixpos) {
  const int mi = blockIdx.x;
  const int n = nn[mi], n_r = nr[mi], n_s = ns[mi], n_bs = nbs[mi];
  const int nss = n_s - n_bs;
  double* Q = R1 + r1_off[mi];
  __shared__ double sh[8];
  __shared__ double P[64 * 64];
  __shared__ double L6[64 * 6];
  __shared__ double dg[64];
  __shared__ int sel[6];
  for (int j = 0; j < 6; j++) {
    for (int i = 0; i < j; i++) {
      double s = 0.0;
      for (int r = threadIdx.x; r < n; r += 256) s += Q[r + (int64_t)j * n] * Q[r + (int64_t)i * n];
      s = block_sum<256>(s, sh);
      for (int r = threadIdx.x; r < n; r += 256) Q[r + (int64_t)j * n] -= Q[r + (int64_t)i * n] * s;
      __syncthreads();
    }
    double s = 0.0;
    for (int r = threadIdx.x; r < n; r += 256) s += Q[r + (int64_t)j * n] * Q[r + (int64_t)j * n];
    s = sqrt(block_sum<256>(s, sh));
    for (int r = threadIdx.x; r < n; r += 256) Q[r + (int64_t)j * n] /= s;
    __syncthreads();
  }
  const int* pv = perm + perm_off[mi];
  double* Rl = R + roff[mi];
  double* B = Rb + rb_off[mi];
  for (int r = threadIdx.x; r < n; r += 256)
    for (int k = 0; k < 6; k++) Rl[pv[r] + (int64_t)k * n] = Q[r + (int64_t)k * n];
  for (int p = threadIdx.x; p < n_s; p += 256)
    for (int k = 0; k < 6; k++) B[p + (int64_t)k * n_s] = Q[(n_r + p) + (int64_t)k * n];
  // P = Rs Rs^T on the s-set (last nss rows of the b-order)
  for (int idx = threadIdx.x; idx < nss * nss; idx += 256) {
    const int a = idx / nss, b = idx % nss;
    double s = 0.0;
    for (int k = 0; k < 6; k++)
      s += Q[(n_r + n_bs + a) + (int64_t)k * n] * Q[(n_r + n_bs + b) + (int64_t)k * n];
    P[a * 64 + b] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < nss; a++) dg[a] = P[a * 64 + a];
    for (int k = 0; k < 6; k++) {
      double mx = -1.0;
      for (int a = 0; a < nss; a++) {
        bool taken = false;
        for (int c = 0; c < k; c++) taken |= (sel[c] == a);
        if (!taken && dg[a] > mx) mx = dg[a];
      }
      int p = -1;
      for (int a = 0; a < nss && p < 0; a++) {
        bool taken = false;
        for (int c = 0; c < k; c++) taken |= (sel[c] == a);
        if (!taken && dg[a] >= mx * (1.0 - 1e-9)) p = a;
      }
      if (p < 0) {  // non-finite kernel basis: flag it, keep indices in range
        fixpos[mi * 6 + k] = -1;
        for (int a = 0; a < nss && p < 0; a++) {
          bool taken = false;
          for (int c = 0; c < k; c++) taken |= (sel[c] == a);
          if (!taken) p = a;
        }
      }
      sel[k] = p;
      const double piv = sqrt(dg[p]);
      for (int a = 0; a < nss; a++) {
        double v = P[a * 64 + p];
        for (int c = 0; c < k; c++) v -= L6[a * 6 + c] * L6[p * 6 + c];
        L6[a * 6 + k] = v / piv;
      }
      for (int a = 0; a < nss; a++) dg[a] -= L6[a * 6 + k] * L6[a * 6 + k];
    }
    // regularise: M2 diagonal at the fixing positions += K_pp (original diagonal)
    double* A2 = M2 + m2off[mi];
    const int64_t l2 = ld2[mi];
    const double* kd = kdiag + kd_off[mi];
    for (int k = 0; k < 6; k++) {
      const int pos = n_bs + sel[k];
      const int ldof = pv[n_r + pos];
      A2[pos + (int64_t)pos * l2] += kd[ldof];
      if (fixpos[mi * 6 + k] != -1) fixpos[mi * 6 + k] = pos;
    }
  }
}

hipError_t launch_gs_pivots(hipStream_t st, int nmat, double* R, const int64_t* roff, const int* n,
                            const int* perm, const int64_t* perm_off, const int* nr, const int* ns,
                            const int* nbs, double* Rb, const int64_t* rb_off, double* Rblk,
                            const int64_t* rblk_off, const double* kdiag, const int64_t* kd_off, double* M2,
                            const int64_t* m2off, const int* ld2, int* fixpos) {
  hipLaunchKernelGGL(k_gs_pivots, dim3(nmat), dim3(256), 0, st, R, roff, n, perm, perm_off, nr, ns, nbs, Rb, rb_off,
                     Rblk, rblk_off, kdiag, kd_off, M2, m2off, ld2, fixpos);
  return hipGetLastError();
}

// ============================================================ PCPG: gather + GEMV
// y_i[rows] = A_i[rows,:] x_i with x_i = B_i^T (wgt .* v) gathered into LDS
// (or x_i = xb_i if xb != null).  A_i symmetric, column major, ld even; each
// thread owns two consecutive rows and streams 16-byte loads down the columns.
__global__ __launch_bounds__(256) void k_gemv_gather(const double* __restrict__ A, const int64_t* __restrict__ aoff,
                                                     const int* __restrict__ lda, const int* __restrict__ ns,
                                                     const int64_t* __restrict__ boff, const int* __restrict__ gptr,
                                                     const int* __restrict__ gidx, const double* __restrict__ v,
                                                     const double* __restrict__ wgt, const double* __restrict__ xb,
                                                     const int64_t* __restrict__ xboff, double* __restrict__ y,
                                                     const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ double xs[];
  const int d = blockIdx.y;
  const int n = ns[d];
  const int r0 = blockIdx.x * 512;
  if (r0 >= n) return;
  const int64_t b0 = boff[d];
  for (int p = threadIdx.x; p < n; p += 256) {
    double s;
    if (xb) {
      s = xb[xboff[d] + p];
    } else {
      s = 0.0;
      const int e1 = gptr[b0 + p + 1];
      for (int e = gptr[b0 + p]; e < e1; e++) {
        const int code = gidx[e];
        const int r = code >= 0 ? code : ~code;
        double val = v[r];
        if (wgt) val *= wgt[r];
        s += code >= 0 ? val : -val;
      }
    }
    xs[p] = s;
  }
  __syncthreads();
  const double* Ad = A + aoff[d];
  const int64_t l = lda[d];
  const int r = r0 + 2 * threadIdx.x;
  if (r >= n) return;
  double a0 = 0.0, a1 = 0.0, b0v = 0.0, b1v = 0.0;
  if (r + 1 < n) {
    const double* col = Ad + r;
    int j = 0;
    for (; j + 4 <= n; j += 4) {
      const double2 c0 = *reinterpret_cast<const double2*>(col + (int64_t)(j + 0) * l);
      const double2 c1 = *reinterpret_cast<const double2*>(col + (int64_t)(j + 1) * l);
      const double2 c2 = *reinterpret_cast<const double2*>(col + (int64_t)(j + 2) * l);
      const double2 c3 = *reinterpret_cast<const double2*>(col + (int64_t)(j + 3) * l);
      const double x0 = xs[j], x1 = xs[j + 1], x2 = xs[j + 2], x3 = xs[j + 3];
      a0 += c0.x * x0; a1 += c0.y * x0;
      b0v += c1.x * x1; b1v += c1.y * x1;
      a0 += c2.x * x2; a1 += c2.y * x2;
      b0v += c3.x * x3; b1v += c3.y * x3;
    }
    for (; j < n; j++) {
      const double2 c0 = *reinterpret_cast<const double2*>(col + (int64_t)j * l);
      a0 += c0.x * xs[j];
      a1 += c0.y * xs[j];
    }
    y[b0 + r] = a0 + b0v;
    y[b0 + r + 1] = a1 + b1v;
  } else {
    for (int j = 0; j < n; j++) a0 += Ad[r + (int64_t)j * l] * xs[j];
    y[b0 + r] = a0;
  }
}

hipError_t launch_gemv_gather(hipStream_t st, int nsub, int maxrows, const double* A, const int64_t* aoff,
                              const int* lda, const int* ns, const int64_t* boff, const int* gptr,
                              const int* gidx, const double* v, const double* wgt, const double* xb,
                              const int64_t* xboff, double* y, const int* done) {
  const int gx = (maxrows + 511) / 512;
  hipLaunchKernelGGL(k_gemv_gather, dim3(gx, nsub), dim3(256), (size_t)maxrows * sizeof(double), st, A, aoff, lda,
                     ns, boff, gptr, gidx, v, wgt, xb, xboff, y, done);
  return hipGetLastError();
}

// ============================================================ cluster correction
// per subdomain: c_i = C_i^T x_i (C_i = K+_bb Bc_i^T), a_i = R_b,i^T x_i.  One
// 1024-thread workgroup per subdomain: wave w owns rq*64 consecutive rows of each
// segment of 1024*rq rows (rq <= 4 rows per lane, held in registers), every lane
// keeps the wave's sum of column k0 + lane over the segments, and the 16 wave sums
// are added in wave order: a fixed reduction order, so the result is reproducible.
// Columns go in passes of 64 (C_i then R_b,i).
static DEV double xb_or_gather(const double* __restrict__ xb, const int* __restrict__ gptr,
                               const int* __restrict__ gidx, const double* __restrict__ v, int64_t q) {
  if (xb) return xb[q];
  double s = 0.0;
  const int e1 = gptr[q + 1];
  for (int e = gptr[q]; e < e1; e++) {
    const int code = gidx[e];
    const int r = code >= 0 ? code : ~code;
    s += code >= 0 ? v[r] : -v[r];
  }
  return s;
}

constexpr int CD_RQ = 4;

__global__ __launch_bounds__(1024) void k_corr_dots(const double* __restrict__ C, const int64_t* __restrict__ coff,
                                                    const int* __restrict__ ldc, const int* __restrict__ nlc,
                                                    const int64_t* __restrict__ lcoff, const double* __restrict__ Rb,
                                                    const int64_t* __restrict__ rboff, const int* __restrict__ ns,
                                                    const int64_t* __restrict__ boff, const int* __restrict__ gptr,
                                                    const int* __restrict__ gidx, const double* __restrict__ v,
                                                    const double* __restrict__ xb, double* __restrict__ cvec,
                                                    double* __restrict__ avec, const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double red[16 * 64];
  const int d = blockIdx.x;
  const int n = ns[d];
  const int64_t b0 = boff[d];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nc = nlc[d], ncol = nc + 6;
  const double* Cd = C + coff[d];
  const int64_t l = ldc[d];
  const double* Rd = Rb + rboff[d];
  const int rq = min(CD_RQ, (n + 1023) / 1024);
  for (int k0 = 0; k0 < ncol; k0 += 64) {
    const int kn = min(64, ncol - k0);
    double mine = 0.0;  // this wave's sum for column k0 + lane
    for (int s0 = 0; s0 < n; s0 += 1024 * rq) {
      double xr[CD_RQ];
      int pr[CD_RQ];
#pragma unroll
      for (int q = 0; q < CD_RQ; q++) {
        pr[q] = s0 + (wv * rq + q) * 64 + lane;
        xr[q] = (q < rq && pr[q] < n) ? xb_or_gather(xb, gptr, gidx, v, b0 + pr[q]) : 0.0;
      }
      // 8 columns at a time: all their loads are issued before the reductions
      for (int kk = 0; kk < kn; kk += 8) {
        double sum[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          sum[u] = 0.0;
          const int k = k0 + kk + u;
          if (kk + u < kn) {
            const double* col = k < nc ? Cd + (int64_t)k * l : Rd + (int64_t)(k - nc) * n;
#pragma unroll
            for (int q = 0; q < CD_RQ; q++)
              if (q < rq && pr[q] < n) sum[u] += col[pr[q]] * xr[q];
          }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const double w = wave_sum(sum[u]);
          if (lane == kk + u) mine += w;
        }
      }
    }
    red[wv * 64 + lane] = mine;
    __syncthreads();
    if ((int)threadIdx.x < kn) {
      double sum = 0.0;
      for (int w = 0; w < 16; w++) sum += red[w * 64 + threadIdx.x];
      const int k = k0 + threadIdx.x;
      if (k < nc) cvec[lcoff[d] + k] = sum;
      else avec[6 * d + (k - nc)] = sum;
    }
    __syncthreads();
  }
}

hipError_t launch_corr_dots(hipStream_t st, int nsub, int maxns, const double* C, const int64_t* coff,
                               const int* ldc, const int* nlc, const int64_t* lcoff, const double* Rb,
                               const int64_t* rboff, const int* ns, const int64_t* boff, const int* gptr,
                               const int* gidx, const double* v, const double* xb, double* cvec, double* avec,
                               const int* done) {
  (void)maxns;
  hipLaunchKernelGGL(k_corr_dots, dim3(nsub), dim3(1024), 0, st, C, coff, ldc, nlc, lcoff, Rb, rboff, ns, boff, gptr,
                     gidx, v, xb, cvec, avec, done);
  return hipGetLastError();
}

// gc = [sum_i scatter(c_i) ; -a_i]  (Cluster.cpp:1129-1153), per cluster
__global__ __launch_bounds__(256) void k_cluster_gc(const int* __restrict__ gc_src, const int64_t* __restrict__ gc_off,
                                                    const int* __restrict__ nlcl, const int* __restrict__ nrcl,
                                                    const int* __restrict__ d0, const double* __restrict__ cvec,
                                                    const double* __restrict__ avec, double* __restrict__ gc,
                                                    const int64_t* __restrict__ la_off,
                                                    const int* __restrict__ done) {
  if (done && *done) return;
  const int cl = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nl = nlcl[cl], nr = nrcl[cl];
  if (i >= nl + nr) return;
  double v;
  if (i < nl) {
    const int* s = gc_src + 2 * (gc_off[cl] + i);
    v = cvec[s[0]];
    if (s[1] >= 0) v += cvec[s[1]];
  } else {
    const int m = i - nl;
    v = -avec[6 * (d0[cl] + m / 6) + m % 6];
  }
  gc[la_off[cl] + i] = v;
}

hipError_t launch_cluster_gc(hipStream_t st, int ncl, int maxlen, const int* gc_src, const int64_t* gc_off,
                             const int* nlcl, const int* nrcl, const int* d0, const double* cvec,
                             const double* avec, double* gc, const int64_t* la_off, const int* done) {
  hipLaunchKernelGGL(k_cluster_gc, dim3((maxlen + 255) / 256, ncl), dim3(256), 0, st, gc_src, gc_off, nlcl, nrcl, d0,
                     cvec, avec, gc, la_off, done);
  return hipGetLastError();
}

// [lam_c; alpha] = Ac^-1 gc with the explicit (symmetric) inverse, per cluster
__global__ __launch_bounds__(256) void k_cluster_gemv(const double* __restrict__ Ainv,
                                                      const int64_t* __restrict__ aoff, const int* __restrict__ nlen,
                                                      const double* __restrict__ gc, const int64_t* __restrict__ la_off,
                                                      double* __restrict__ la, const int* __restrict__ done) {
  if (done && *done) return;
  extern __shared__ double xs[];
  const int cl = blockIdx.y;
  const int n = nlen[cl];
  const int r = blockIdx.x * 256 + threadIdx.x;
  if ((int)(blockIdx.x * 256) >= n) return;
  const int64_t o = la_off[cl];
  for (int p = threadIdx.x; p < n; p += 256) xs[p] = gc[o + p];
  __syncthreads();
  if (r >= n) return;
  const double* A = Ainv + aoff[cl];
  double s = 0.0;
  for (int j = 0; j < n; j++) s += A[r + (int64_t)j * n] * xs[j];
  la[o + r] = s;
}

hipError_t launch_cluster_gemv(hipStream_t st, int ncl, int maxlen, const double* Ainv, const int64_t* aoff,
                               const int* nlen, const double* gc, const int64_t* la_off, double* la,
                               const int* done) {
  hipLaunchKernelGGL(k_cluster_gemv, dim3((maxlen + 255) / 256, ncl), dim3(256), (size_t)maxlen * sizeof(double), st,
                     Ainv, aoff, nlen, gc, la_off, la, done);
  return hipGetLastError();
}

// la[o + i] = sum_j A[j + i*n] gc[o + j], A symmetric (column i = row i): one
// wave per output, lanes over j (coalesced), fixed butterfly reduction
__global__ __launch_bounds__(256) void k_sym_gemv_waves(const double* __restrict__ A,
                                                        const int64_t* __restrict__ aoff,
                                                        const int* __restrict__ nlen, const double* __restrict__ x,
                                                        const int64_t* __restrict__ xoff, double* __restrict__ out,
                                                        const int* __restrict__ done) {
  if (done && *done) return;
  const int cl = blockIdx.y;
  const int n = nlen[cl];
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  const double* col = A + aoff[cl] + (int64_t)i * n;
  const double* xv = x + xoff[cl];
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += col[j] * xv[j];
  s = wave_sum(s);
  if (lane == 0) out[xoff[cl] + i] = s;
}

hipError_t launch_sym_gemv_waves(hipStream_t st, int ncl, int maxlen, const double* A, const int64_t* aoff,
                                 const int* nlen, const double* x, const int64_t* xoff, double* out,
                                 const int* done) {
  hipLaunchKernelGGL(k_sym_gemv_waves, dim3((maxlen + 3) / 4, ncl), dim3(256), 0, st, A, aoff, nlen, x, xoff, out,
                     done);
  return hipGetLastError();
}

// y_i += R_b,i alpha_i - C_i lam_c,i   (Cluster.cpp:1162-1181)
__global__ __launch_bounds__(256) void k_corr_apply(const double* __restrict__ C, const int64_t* __restrict__ coff,
                                                    const int* __restrict__ ldc, const int* __restrict__ nlc,
                                                    const int64_t* __restrict__ lcoff, const int* __restrict__ lcmap,
                                                    const double* __restrict__ Rb, const int64_t* __restrict__ rboff,
                                                    const int* __restrict__ ns, const int64_t* __restrict__ boff,
                                                    const int64_t* __restrict__ alpha_idx,
                                                    const double* __restrict__ la, double* __restrict__ y,
                                                    const int* __restrict__ done) {
  if (done && *done) return;
  const int d = blockIdx.y;
  const int n = ns[d];
  __shared__ double lc[256];
  __shared__ double al[6];
  const int nc = nlc[d];
  if ((int)(blockIdx.x * 256) >= n) return;
  for (int k = threadIdx.x; k < nc; k += 256) lc[k] = la[lcmap[lcoff[d] + k]];
  if (threadIdx.x < 6) al[threadIdx.x] = alpha_idx ? la[alpha_idx[d] + threadIdx.x] : 0.0;
  __syncthreads();
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double* Rd = Rb + rboff[d];
  double s = 0.0;
  for (int k = 0; k < 6; k++) s += Rd[p + (int64_t)k * n] * al[k];
  const double* Cd = C + coff[d];
  const int64_t l = ldc[d];
  double t = 0.0;
  for (int k = 0; k < nc; k++) t += Cd[p + (int64_t)k * l] * lc[k];
  y[boff[d] + p] += s - t;
}

hipError_t launch_corr_apply(hipStream_t st, int nsub, int maxrows, const double* C, const int64_t* coff,
                             const int* ldc, const int* nlc, const int64_t* lcoff, const int* lcmap,
                             const double* Rb, const int64_t* rboff, const int* ns, const int64_t* boff,
                             const int64_t* alpha_idx, const double* la, double* y, const int* done) {
  hipLaunchKernelGGL(k_corr_apply, dim3((maxrows + 255) / 256, nsub), dim3(256), 0, st, C, coff, ldc, nlc, lcoff,
                     lcmap, Rb, rboff, ns, boff, alpha_idx, la, y, done);
  return hipGetLastError();
}

// ============================================================ dual-space kernels
// out[r] = wgt[r] * (y[ra[r]] - y[rb[r]])  (mult_Bf, Cluster.cpp:1193-1208; scale)
// and, if dotw, block partial of sum out[r]*dotw[r].
__global__ __launch_bounds__(256) void k_scatter(int64_t nlam, const int* __restrict__ ra, const int* __restrict__ rb,
                                                 const double* __restrict__ y, const double* __restrict__ wgt,
                                                 const double* __restrict__ sub_from, double* __restrict__ out,
                                                 const double* __restrict__ dotw, double* __restrict__ part,
                                                 const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double sh[8];
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double prod = 0.0;
  if (r < nlam) {
    double v = y[ra[r]];
    const int b = rb[r];
    if (b >= 0) v -= y[b];
    if (wgt) v *= wgt[r];
    if (sub_from) v -= sub_from[r];
    out[r] = v;
    if (dotw) prod = v * dotw[r];
  }
  if (dotw) {
    const double s = block_sum<256>(prod, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

hipError_t launch_scatter(hipStream_t st, int64_t nlam, const int* ra, const int* rb, const double* y,
                          const double* wgt, const double* sub_from, double* out, const double* dotw,
                          double* part, const int* done) {
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((nlam + 255) / 256)), dim3(256), 0, st, nlam, ra, rb, y, wgt,
                     sub_from, out, dotw, part, done);
  return hipGetLastError();
}

// mode 0: wFw -> rho; 1: gPz -> gamma, iteration count and stopping rule
// (Cluster.cpp:1425); 2: initial gPz -> norm0; 3: plain sum into sc->gPg0.
__global__ __launch_bounds__(256) void k_reduce(int nparts, const double* __restrict__ part, Scalars* sc, int mode) {
  if (mode != 2 && mode != 3 && sc->done) return;
  __shared__ double sh[8];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) s += part[i];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      sc->wFw = s;
      sc->rho = -sc->gPz / s;
    } else if (mode == 1) {
      sc->gPz_prev = sc->gPz;
      sc->gPz = s;
      sc->gamma = s / sc->gPz_prev;
      sc->it += 1;
      if (!(sqrt(fabs(s)) >= sc->eps * sc->norm0) || sc->it >= sc->max_iter) sc->done = 1;
      if (!isfinite(s)) {  // stop on a non-finite residual; the host reports the error
        sc->done = 1;
        sc->err = 1;
      }
    } else if (mode == 2) {
      sc->gPz = s;
      sc->norm0 = sqrt(s);
      sc->it = 0;
      // stopping test at the top of the loop (Cluster.cpp:1425); a zero initial
      // residual (zero load) is already the solution: no iteration
      sc->done = (sc->max_iter <= 0 || !(s > 0.0) || !(sqrt(s) >= sc->eps * sc->norm0)) ? 1 : 0;
      if (!isfinite(s)) {
        sc->done = 1;
        sc->err = 1;
      }
    } else {
      sc->gPg0 = s;
    }
  }
}

hipError_t launch_reduce(hipStream_t st, int nparts, const double* part, Scalars* sc, int mode) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, nparts, part, sc, mode);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_axpy2(int64_t nlam, double* __restrict__ lam, double* __restrict__ g,
                                               const double* __restrict__ w, const double* __restrict__ Fw,
                                               const Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= nlam) return;
  const double rho = sc->rho;
  lam[r] += rho * w[r];
  g[r] += rho * Fw[r];
}

hipError_t launch_axpy2(hipStream_t st, int64_t nlam, double* lam, double* g, const double* w, const double* Fw,
                        const Scalars* sc) {
  hipLaunchKernelGGL(k_axpy2, dim3((unsigned)((nlam + 255) / 256)), dim3(256), 0, st, nlam, lam, g, w, Fw, sc);
  return hipGetLastError();
}

// v_i = R_b,i^T (B_i^T x): one 1024-thread workgroup per subdomain, rows strided
// over the threads, the six sums reduced per wave and then in wave order (fixed)
__global__ __launch_bounds__(1024) void k_proj_dots(const double* __restrict__ Rb, const int64_t* __restrict__ rboff,
                                                    const int* __restrict__ ns, const int64_t* __restrict__ boff,
                                                    const int* __restrict__ gptr, const int* __restrict__ gidx,
                                                    const double* __restrict__ x, double* __restrict__ v,
                                                    const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double red[16 * 6];
  const int d = blockIdx.x;
  const int n = ns[d];
  const int64_t b0 = boff[d];
  const double* Rd = Rb + rboff[d];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int p = threadIdx.x; p < n; p += 1024) {
    const double s = xb_or_gather(nullptr, gptr, gidx, x, b0 + p);
#pragma unroll
    for (int k = 0; k < 6; k++) acc[k] += Rd[p + (int64_t)k * n] * s;
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) red[wv * 6 + k] = s;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int w = 0; w < 16; w++) s += red[w * 6 + threadIdx.x];
    v[6 * d + threadIdx.x] = s;
  }
}

hipError_t launch_proj_dots(hipStream_t st, int nsub, const double* Rb, const int64_t* rboff, const int* ns,
                            const int64_t* boff, const int* gptr, const int* gidx, const double* x, double* v,
                            const int* done) {
  hipLaunchKernelGGL(k_proj_dots, dim3(nsub), dim3(1024), 0, st, Rb, rboff, ns, boff, gptr, gidx, x, v, done);
  return hipGetLastError();
}

// y = Gf^T x = [sum_{i in cl} H_i^T v_i]_cl ; a = (Gf^T Gf)^-1 y ; alpha = asgn * a ;
// h_i = sgn * H_i a_cl(i).  One workgroup; subdomains of a cluster are the
// contiguous ids [cl*nsub_cl, (cl+1)*nsub_cl).  (Cluster.cpp:1314-1329)
// The ncol sums of y and of a are each split into P = 1024 / ncol contiguous parts
// (thread c + ncol*part), added part by part in a fixed order.
__global__ __launch_bounds__(1024) void k_proj_coarse(int nsub, int ncol, const double* __restrict__ v,
                                                      const double* __restrict__ H, const int* __restrict__ cl,
                                                      const double* __restrict__ invGtG, double* __restrict__ h,
                                                      double* __restrict__ alpha, double sgn,
                                                      const int* __restrict__ done, double asgn) {
  if (done && *done) return;
  __shared__ double y[1024];
  __shared__ double al[1024];
  __shared__ double red[1024];
  const int ncl = ncol / 6;
  const int nsc = nsub / ncl;
  const int P = 1024 / ncol;
  const int t = threadIdx.x;
  const int c = t % ncol, part = t / ncol;
  // y[c] = sum over the cluster's subdomains of (H_d^T v_d)[k]
  if (part < P) {
    const int ccl = c / 6, k = c % 6;
    const int per = (nsc + P - 1) / P;
    const int d0 = ccl * nsc + part * per, d1 = min(ccl * nsc + (part + 1) * per, (ccl + 1) * nsc);
    double s = 0.0;
    for (int d = d0; d < d1; d++) {
      const double* Hd = H + 36 * d;
      const double* vd = v + 6 * d;
      s += Hd[0 * 6 + k] * vd[0] + Hd[1 * 6 + k] * vd[1] + Hd[2 * 6 + k] * vd[2] + Hd[3 * 6 + k] * vd[3] +
           Hd[4 * 6 + k] * vd[4] + Hd[5 * 6 + k] * vd[5];
    }
    red[t] = s;
  }
  __syncthreads();
  if (t < ncol) {
    double s = 0.0;
    for (int q = 0; q < P; q++) s += red[t + q * ncol];
    y[t] = s;
  }
  __syncthreads();
  // a = (Gf^T Gf)^-1 y
  if (part < P) {
    const int per = (ncol + P - 1) / P;
    const int j0 = part * per, j1 = min(ncol, (part + 1) * per);
    double s = 0.0;
    for (int j = j0; j < j1; j++) s += invGtG[c + (int64_t)j * ncol] * y[j];
    red[t] = s;
  }
  __syncthreads();
  if (t < ncol) {
    double s = 0.0;
    for (int q = 0; q < P; q++) s += red[t + q * ncol];
    al[t] = s;
    if (alpha) alpha[t] = asgn * s;
  }
  __syncthreads();
  for (int q = t; q < 6 * nsub; q += blockDim.x) {
    const int d = q / 6, a = q % 6;
    const double* Hd = H + 36 * d + a * 6;
    const double* ad = al + 6 * cl[d];
    h[q] = sgn * (Hd[0] * ad[0] + Hd[1] * ad[1] + Hd[2] * ad[2] + Hd[3] * ad[3] + Hd[4] * ad[4] + Hd[5] * ad[5]);
  }
}

hipError_t launch_proj_coarse(hipStream_t st, int nsub, int ncol, const double* v, const double* H, const int* cl,
                              const double* invGtG, double* h, double* alpha, double sgn, const int* done,
                              double asgn) {
  if (ncol > 1024 || ncol < 6) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_proj_coarse, dim3(1), dim3(1024), 0, st, nsub, ncol, v, H, cl, invGtG, h, alpha, sgn, done,
                     asgn);
  return hipGetLastError();
}

// out[r] = x[r] + sgn * (Gf h)[r] with (Gf h)[r] = R_b,a[pos_a].h_a - R_b,b[pos_b].h_b ;
// optional block partial of out.dotv
__global__ __launch_bounds__(256) void k_gh(int64_t nlam, const int* __restrict__ rsa, const int* __restrict__ rsb,
                                            const int* __restrict__ rpa, const int* __restrict__ rpb,
                                            const double* __restrict__ Rb, const int64_t* __restrict__ rboff,
                                            const int* __restrict__ ns, const double* __restrict__ h,
                                            const double* __restrict__ x, double* __restrict__ out, double sgn,
                                            const double* __restrict__ dotv, double* __restrict__ part,
                                            const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double sh[8];
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double prod = 0.0;
  if (r < nlam) {
    const int a = rsa[r];
    const int na = ns[a];
    const double* Ra = Rb + rboff[a] + rpa[r];
    double g = 0.0;
    for (int k = 0; k < 6; k++) g += Ra[(int64_t)k * na] * h[6 * a + k];
    const int b = rsb[r];
    if (b >= 0) {
      const int nb = ns[b];
      const double* Rbb = Rb + rboff[b] + rpb[r];
      double gb = 0.0;
      for (int k = 0; k < 6; k++) gb += Rbb[(int64_t)k * nb] * h[6 * b + k];
      g -= gb;
    }
    const double o = (x ? x[r] : 0.0) + sgn * g;
    out[r] = o;
    if (dotv) prod = o * dotv[r];
  }
  if (dotv) {
    const double s = block_sum<256>(prod, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

hipError_t launch_gh(hipStream_t st, int64_t nlam, const int* rsub_a,
                     const int* rsub_b, const int* rpos_a, const int* rpos_b, const double* Rb,
                     const int64_t* rboff, const int* ns, const double* h, const double* x, double* out,
                     double sgn, const double* dotv, double* part, const int* done) {
  hipLaunchKernelGGL(k_gh, dim3((unsigned)((nlam + 255) / 256)), dim3(256), 0, st, nlam, rsub_a, rsub_b, rpos_a,
                     rpos_b, Rb, rboff, ns, h, x, out, sgn, dotv, part, done);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_update_w(int64_t nlam, double* __restrict__ w,
                                                  const double* __restrict__ Pz, const Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= nlam) return;
  w[r] = Pz[r] + sc->gamma * w[r];
}

hipError_t launch_update_w(hipStream_t st, int64_t nlam, double* w, const double* Pz, const Scalars* sc) {
  hipLaunchKernelGGL(k_update_w, dim3((unsigned)((nlam + 255) / 256)), dim3(256), 0, st, nlam, w, Pz, sc);
  return hipGetLastError();
}

}  // namespace hfls
