// Interior solves with the Cholesky factor L of K_rr (the interior block of a
// subdomain stiffness matrix, interior-first order): K_rr^-1 v = W^T W v with
// W = L^-1, kept explicitly so that both products are parallel tile products
// instead of sequential triangular solves.
//
// L and W are packed lower 64x64 tiles (tile (I,J), I >= J, at
// J*T - J*(J-1)/2 + (I-J), column major inside a tile, zero padded).  During the
// setup the diagonal tiles of L are replaced by their inverses D_I = L_II^-1 and W
// is formed from them block column by block column.  The coupling K_rb (b rows or
// interior rows) is kept sparse (CSR over packed rows, values taken from the
// assembled matrix): its stencil has at most 81 entries per row.
//
// Vector blocks: entry (i, j) of subdomain d's block (row i < n_r, column j < k)
// at X[xbase[d] + i*rs + j*cs], so one kernel serves row-major blocks of k load
// cases (rs = 64, cs = 1), single packed vectors (k = 1) and column-major n x 6
// bases (rs = 1, cs = n; a per-subdomain cs array overrides the scalar).
// (csr_apply: csv, when given, is the column stride of both X and Y.)
#include "kernels.hpp"

namespace hfls {

typedef double lrr_d4 __attribute__((ext_vector_type(4)));

static __device__ __forceinline__ int64_t ltile(int T, int I, int J) {
  return (int64_t)J * T - (int64_t)J * (J - 1) / 2 + (I - J);
}

// ------------------------------------------------------------ diagonal inverses
// One 64-thread workgroup per (diagonal tile I, subdomain): the w x w lower block
// (w = min(64, n_r - 64 I)) is replaced by its inverse (forward substitution, one
// column per thread); everything outside w x w becomes 0.
__global__ __launch_bounds__(64) void k_lrr_diag_inv(const int* __restrict__ nr, double* __restrict__ Lt,
                                                     const int64_t* __restrict__ loff) {
  const int d = blockIdx.y, I = blockIdx.x, t = threadIdx.x;
  const int n = nr[d], T = (n + 63) / 64;
  if (I >= T) return;
  const int w = min(64, n - 64 * I);
  double* D = Lt + loff[d] + ltile(T, I, I) * 4096;
  __shared__ double L[64 * 65];
  __shared__ double Y[64 * 65];
  for (int c = 0; c < 64; c++) L[c * 65 + t] = D[c * 64 + t];
  __syncthreads();
  for (int i = 0; i < 64; i++) {  // column t of the inverse
    double y = 0.0;
    if (t < w && i < w && i >= t) {
      double s = i == t ? 1.0 : 0.0;
      for (int c = t; c < i; c++) s -= L[c * 65 + i] * Y[t * 65 + c];
      y = s / L[i * 65 + i];
    }
    Y[t * 65 + i] = y;
  }
  __syncthreads();
  for (int c = 0; c < 64; c++) D[c * 64 + t] = Y[c * 65 + t];
}

hipError_t launch_lrr_diag_inv(hipStream_t st, int nsub, int maxT, const int* nr, double* Lt, const int64_t* loff) {
  if (maxT < 1) return hipSuccess;
  hipLaunchKernelGGL(k_lrr_diag_inv, dim3(maxT, nsub), dim3(64), 0, st, nr, Lt, loff);
  return hipGetLastError();
}

// ------------------------------------------------------------ inverse of L
// W = L^-1 (lower), column block J by workgroup (J, subdomain): for I = J..T-1,
//   W_IJ = D_I (delta_IJ 1 - sum_{J<=K<I} L_IK W_KJ),
// where the diagonal tiles of L already hold D_I = L_II^-1.  Each product is a
// 64x64 tile pair on the matrix cores (v_mfma_f64_16x16x4f64; wave v owns rows
// 16v..16v+15 and the four 16-column blocks; lane layouts A[l%16][l/16],
// B[l/16][l%16], D row 4r+l/16, col l%16); the next pair is loaded into registers
// while the current one is multiplied.  W is written in the packed tile layout of
// L; the block column's own tiles are read back by the same workgroup.
struct InvStage {
  double a[16], b[16];
};

__global__ __launch_bounds__(256) void k_lrr_trtri(const int* __restrict__ nr, const double* __restrict__ Lt,
                                                   const int64_t* __restrict__ loff, double* __restrict__ Wt,
                                                   const int64_t* __restrict__ woff) {
  const int d = blockIdx.y, J = blockIdx.x;
  const int n = nr[d], T = (n + 63) / 64;
  if (J >= T) return;
  const double* L = Lt + loff[d];
  double* W = Wt + woff[d];
  __shared__ double As[64 * 65];  // [kk][row]
  __shared__ double Bs[64 * 65];  // [kk][col]
  const int t = threadIdx.x, lane = t & 63, v = t >> 6;
  // A = L tile (column major: A(r, kk) = tile[kk*64 + r]); B = W tile (B(kk, c) = tile[c*64 + kk])
  auto fetch = [&](const double* at, const double* bt, InvStage& s) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = t + 256 * q;
      s.a[q] = at[e];
      s.b[q] = bt[e];
    }
  };
  auto stash = [&](const InvStage& s) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int e = t + 256 * q;
      const int lo = e & 63, hi = e >> 6;
      As[hi * 65 + lo] = s.a[q];  // A(row lo, kk hi)
      Bs[lo * 65 + hi] = s.b[q];  // B(kk lo, col hi)
    }
  };
  auto mma = [&](lrr_d4* acc) {
    for (int kk = 0; kk < 64; kk += 4) {
      const double a = As[(kk + (lane >> 4)) * 65 + 16 * v + (lane & 15)];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double b = Bs[(kk + (lane >> 4)) * 65 + 16 * c + (lane & 15)];
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
      }
    }
  };
  for (int I = J; I < T; I++) {
    lrr_d4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[c][r] = 0.0;
    const int nK = I - J;  // K = J..I-1
    InvStage s;
    if (nK > 0) fetch(L + ltile(T, I, J) * 4096, W + ltile(T, J, J) * 4096, s);
    for (int q = 0; q < nK; q++) {
      __syncthreads();
      stash(s);
      __syncthreads();
      if (q + 1 < nK) fetch(L + ltile(T, I, J + q + 1) * 4096, W + ltile(T, J + q + 1, J) * 4096, s);
      mma(acc);
    }
    // R = delta_IJ 1 - acc into Bs ([row][col] = B(kk = row, col)); D_I into As
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 16 * v + 4 * r + (lane >> 4), col = 16 * c + (lane & 15);
        const double e = (I == J && row == col && 64 * I + row < n) ? 1.0 : 0.0;
        Bs[row * 65 + col] = e - acc[c][r];
      }
    {
      const double* D = L + ltile(T, I, I) * 4096;
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int e = t + 256 * q;
        As[(e >> 6) * 65 + (e & 63)] = D[e];
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[c][r] = 0.0;
    mma(acc);
    double* out = W + ltile(T, I, J) * 4096;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 16 * v + 4 * r + (lane >> 4), col = 16 * c + (lane & 15);
        out[col * 64 + row] = acc[c][r];
      }
    // the next block rows read this tile back
    __threadfence_block();
    __syncthreads();
  }
}

hipError_t launch_lrr_trtri(hipStream_t st, int nsub, int maxT, const int* nr, const double* Lt, const int64_t* loff,
                            double* Wt, const int64_t* woff) {
  if (maxT < 1 || nsub < 1) return hipSuccess;
  hipLaunchKernelGGL(k_lrr_trtri, dim3(maxT, nsub), dim3(256), 0, st, nr, Lt, loff, Wt, woff);
  return hipGetLastError();
}

// ------------------------------------------------------------ products with W = L^-1
// k columns of an interior block (element (i, j) of subdomain d at
// X[xbase[d] + i*rs + j*cs]); op 0: Y = W X, op 1: Y = W^T X.  One 256-thread
// workgroup per (row block I, subdomain); every tile is read once per column.
// op 0 (Y_I = sum_{J<=I} W_IJ X_J): lane = row; wave w takes the tiles J = w, w+4,
// .. whole (64 columns each), two at a time, so each wave streams its own tiles
// with 32 loads in flight; the 4 wave sums are added in LDS.  op 1 (Y_I = sum_{J>=I}
// W_JI^T X_J): wave w takes rows 16w..16w+15 of the result, lane = tile row c, two
// tiles per step; each wave sums its 16 values over the lanes at the end.  Fixed
// order throughout.
__global__ __launch_bounds__(256) void k_trmv(int op, int k, const int* __restrict__ nr, const double* __restrict__ Wt,
                                              const int64_t* __restrict__ woff, const double* __restrict__ X,
                                              const int64_t* __restrict__ xbase, int64_t rs, int64_t cs0,
                                              double* __restrict__ Y, const int64_t* __restrict__ ybase,
                                              const int* __restrict__ csv) {
  const int d = blockIdx.y, I = blockIdx.x;
  const int n = nr[d], T = (n + 63) / 64;
  if (I >= T) return;
  const int64_t cs = csv ? csv[d] : cs0;
  const double* W = Wt + woff[d];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  __shared__ double red[4][64];
  __shared__ double xs[64 * 64];  // op 0: the x blocks of the row block's tiles, 64 at a time
  for (int j = 0; j < k; j++) {
    const double* x = X + xbase[d] + j * cs;
    double* y = Y + ybase[d] + j * cs;
    if (op == 0) {
      double acc = 0.0;
      for (int J0 = 0; J0 <= I; J0 += 64) {
        const int nJ = min(64, I + 1 - J0);
        __syncthreads();
        for (int e = t; e < nJ * 64; e += 256) {
          const int gc = 64 * J0 + e;
          xs[e] = gc < n ? x[gc * rs] : 0.0;
        }
        __syncthreads();
        int q = w;
        for (; q + 4 < nJ; q += 8) {
          const double* t0 = W + ltile(T, I, J0 + q) * 4096 + lane;
          const double* t1 = W + ltile(T, I, J0 + q + 4) * 4096 + lane;
          double a0[64], a1[64];
#pragma unroll
          for (int c = 0; c < 64; c++) a0[c] = t0[c * 64];
#pragma unroll
          for (int c = 0; c < 64; c++) a1[c] = t1[c * 64];
#pragma unroll
          for (int c = 0; c < 64; c++) acc += a0[c] * xs[q * 64 + c];
#pragma unroll
          for (int c = 0; c < 64; c++) acc += a1[c] * xs[(q + 4) * 64 + c];
        }
        for (; q < nJ; q += 4) {
          const double* t0 = W + ltile(T, I, J0 + q) * 4096 + lane;
          double a0[64];
#pragma unroll
          for (int c = 0; c < 64; c++) a0[c] = t0[c * 64];
#pragma unroll
          for (int c = 0; c < 64; c++) acc += a0[c] * xs[q * 64 + c];
        }
      }
      red[w][lane] = acc;
      __syncthreads();
      if (w == 0 && 64 * I + lane < n)
        y[(64 * I + lane) * rs] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
      __syncthreads();
    } else {
      double acc[16];
#pragma unroll
      for (int m = 0; m < 16; m++) acc[m] = 0.0;
      int J = I;
      for (; J + 1 < T; J += 2) {
        const double* t0 = W + ltile(T, J, I) * 4096 + 16 * w * 64 + lane;
        const double* t1 = W + ltile(T, J + 1, I) * 4096 + 16 * w * 64 + lane;
        const int g0 = 64 * J + lane, g1 = g0 + 64;
        const double x0 = g0 < n ? x[g0 * rs] : 0.0, x1 = g1 < n ? x[g1 * rs] : 0.0;
        double a0[16], a1[16];
#pragma unroll
        for (int m = 0; m < 16; m++) a0[m] = t0[m * 64];
#pragma unroll
        for (int m = 0; m < 16; m++) a1[m] = t1[m * 64];
#pragma unroll
        for (int m = 0; m < 16; m++) acc[m] += a0[m] * x0;
#pragma unroll
        for (int m = 0; m < 16; m++) acc[m] += a1[m] * x1;
      }
      if (J < T) {
        const double* t0 = W + ltile(T, J, I) * 4096 + 16 * w * 64 + lane;
        const int g0 = 64 * J + lane;
        const double x0 = g0 < n ? x[g0 * rs] : 0.0;
#pragma unroll
        for (int m = 0; m < 16; m++) acc[m] += t0[m * 64] * x0;
      }
#pragma unroll
      for (int m = 0; m < 16; m++) {
        double v = acc[m];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        acc[m] = v;
      }
      if (lane < 16) {
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < 16; m++)
          if (lane == m) v = acc[m];
        const int row = 64 * I + 16 * w + lane;
        if (row < n) y[row * rs] = v;
      }
    }
  }
}

hipError_t launch_trmv(hipStream_t st, int nsub, int maxT, int op, int k, const int* nr, const double* Wt,
                       const int64_t* woff, const double* X, const int64_t* xbase, int64_t rs, int64_t cs,
                       double* Y, const int64_t* ybase, const int* csv) {
  if (maxT < 1 || nsub < 1) return hipSuccess;
  if (k < 1 || (op != 0 && op != 1)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_trmv, dim3(maxT, nsub), dim3(256), 0, st, op, k, nr, Wt, woff, X, xbase, rs, cs, Y, ybase,
                     csv);
  return hipGetLastError();
}

// ------------------------------------------------------------ sparse coupling
// Y = beta*Y + alpha * K X over CSR rows [row0, row1): row r belongs to subdomain
// rsub[r] (index into the base arrays after subtracting sub0), local row rloc[r];
// entry e couples it to local column col[e].  One wave per row, lanes = columns.
// k == 1: one thread per row
__global__ __launch_bounds__(256) void k_csr_apply1(int64_t row0, int64_t row1, int sub0,
                                                    const int64_t* __restrict__ ptr, const int* __restrict__ col,
                                                    const double* __restrict__ val, const int* __restrict__ rsub,
                                                    const int* __restrict__ rloc, const double* __restrict__ X,
                                                    const int64_t* __restrict__ xbase, int64_t xrs,
                                                    double* __restrict__ Y, const int64_t* __restrict__ ybase,
                                                    int64_t yrs, double alpha, double beta) {
  const int64_t r = row0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= row1) return;
  const int s = rsub[r] - sub0;
  const double* Xs = X + xbase[s];
  double acc = 0.0;
  for (int64_t e = ptr[r]; e < ptr[r + 1]; e++) acc += val[e] * Xs[(int64_t)col[e] * xrs];
  double* y = Y + ybase[s] + (int64_t)rloc[r] * yrs;
  *y = (beta == 0.0 ? 0.0 : beta * *y) + alpha * acc;
}

__global__ __launch_bounds__(256) void k_csr_apply(int64_t row0, int64_t row1, int sub0, int k,
                                                   const int64_t* __restrict__ ptr, const int* __restrict__ col,
                                                   const double* __restrict__ val, const int* __restrict__ rsub,
                                                   const int* __restrict__ rloc, const double* __restrict__ X,
                                                   const int64_t* __restrict__ xbase, int64_t xrs, int64_t xcs0,
                                                   double* __restrict__ Y, const int64_t* __restrict__ ybase,
                                                   int64_t yrs, int64_t ycs0, const int* __restrict__ csv,
                                                   double alpha, double beta) {
  const int j = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t r = row0 + (int64_t)blockIdx.x * 4 + wv; r < row1; r += nw) {
    const int s = rsub[r] - sub0;
    const int64_t xcs = csv ? csv[s] : xcs0, ycs = csv ? csv[s] : ycs0;
    const double* Xs = X + xbase[s] + j * xcs;
    double acc = 0.0;
    if (j < k)
      for (int64_t e = ptr[r]; e < ptr[r + 1]; e++) acc += val[e] * Xs[(int64_t)col[e] * xrs];
    if (j < k) {
      double* y = Y + ybase[s] + (int64_t)rloc[r] * yrs + j * ycs;
      *y = (beta == 0.0 ? 0.0 : beta * *y) + alpha * acc;
    }
  }
}

hipError_t launch_csr_apply(hipStream_t st, int64_t row0, int64_t row1, int sub0, int k, const int64_t* ptr,
                            const int* col, const double* val, const int* rsub, const int* rloc, const double* X,
                            const int64_t* xbase, int64_t xrs, int64_t xcs, double* Y, const int64_t* ybase,
                            int64_t yrs, int64_t ycs, const int* csv, double alpha, double beta) {
  if (k < 1 || k > 64) return hipErrorInvalidValue;
  if (row1 <= row0) return hipSuccess;
  if (k == 1) {
    hipLaunchKernelGGL(k_csr_apply1, dim3((unsigned)((row1 - row0 + 255) / 256)), dim3(256), 0, st, row0, row1, sub0,
                       ptr, col, val, rsub, rloc, X, xbase, xrs, Y, ybase, yrs, alpha, beta);
    return hipGetLastError();
  }
  const int64_t g = std::min<int64_t>((row1 - row0 + 3) / 4, 4096);
  hipLaunchKernelGGL(k_csr_apply, dim3((unsigned)g), dim3(256), 0, st, row0, row1, sub0, k, ptr, col, val, rsub, rloc,
                     X, xbase, xrs, xcs, Y, ybase, yrs, ycs, csv, alpha, beta);
  return hipGetLastError();
}

// CSR values from the assembled dense lower matrices M (interior-first order, the
// b rows after the n_r interior ones): entries [e0, e1) of rows in subdomains
// sub0.. of this chunk.  trans 0: rows are b positions p, columns interior q,
// value M(n_r + p, q); trans 1: rows interior q, columns b positions p, same value.
__global__ __launch_bounds__(256) void k_csr_extract(int64_t row0, int64_t row1, int sub0, int trans,
                                                     const int64_t* __restrict__ ptr, const int* __restrict__ col,
                                                     const int* __restrict__ rsub, const int* __restrict__ rloc,
                                                     const int* __restrict__ nr, const double* __restrict__ M,
                                                     const int64_t* __restrict__ moff, const int* __restrict__ ld,
                                                     double* __restrict__ val) {
  const int64_t r = row0 + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= row1) return;
  const int s = rsub[r] - sub0;
  const double* Ms = M + moff[s];
  const int64_t l = ld[s];
  for (int64_t e = ptr[r] + (threadIdx.x & 63); e < ptr[r + 1]; e += 64) {
    const int p = trans ? col[e] : rloc[r], q = trans ? rloc[r] : col[e];
    val[e] = Ms[(int64_t)(nr[s] + p) + (int64_t)q * l];
  }
}

hipError_t launch_csr_extract(hipStream_t st, int64_t row0, int64_t row1, int sub0, int trans, const int64_t* ptr,
                              const int* col, const int* rsub, const int* rloc, const int* nr, const double* M,
                              const int64_t* moff, const int* ld, double* val) {
  if (row1 <= row0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_extract, dim3((unsigned)((row1 - row0 + 3) / 4)), dim3(256), 0, st, row0, row1, sub0,
                     trans, ptr, col, rsub, rloc, nr, M, moff, ld, val);
  return hipGetLastError();
}

// B = alpha*A + beta*B on the first n_r rows of every subdomain's block
__global__ __launch_bounds__(256) void k_rr_axpby(int k, const int* __restrict__ nr, const double* __restrict__ A,
                                                  const int64_t* __restrict__ abase, double* __restrict__ B,
                                                  const int64_t* __restrict__ bbase, int64_t rs, int64_t cs,
                                                  double alpha, double beta) {
  const int d = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t i = e / k, j = e % k;
  if (i >= nr[d]) return;
  const int64_t o = i * rs + j * cs;
  B[bbase[d] + o] = alpha * A[abase[d] + o] + beta * B[bbase[d] + o];
}

hipError_t launch_rr_axpby(hipStream_t st, int nsub, int maxnr, int k, const int* nr, const double* A,
                           const int64_t* abase, double* B, const int64_t* bbase, int64_t rs, int64_t cs,
                           double alpha, double beta) {
  if (maxnr < 1 || nsub < 1) return hipSuccess;
  hipLaunchKernelGGL(k_rr_axpby, dim3((unsigned)(((int64_t)maxnr * k + 255) / 256), nsub), dim3(256), 0, st, k, nr,
                     A, abase, B, bbase, rs, cs, alpha, beta);
  return hipGetLastError();
}

}  // namespace hfls
