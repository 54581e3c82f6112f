// Helper kernels of the HTFETI path: permutations between the local DOF order
// and the interior-first order, block copies, small dense products and the
// recovery scatter.  All per-subdomain batched; one launch covers every
// subdomain of a mesh.
#include "kernels.hpp"

namespace hfls {

static __device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// dst[doff+i] = src[soff + perm[poff+i]]  (gather) or the inverse scatter
__global__ __launch_bounds__(256) void k_permute(int scatter, const int* __restrict__ len,
                                                 const int* __restrict__ perm, const int64_t* __restrict__ poff,
                                                 const double* __restrict__ src, const int64_t* __restrict__ soff,
                                                 double* __restrict__ dst, const int64_t* __restrict__ doff,
                                                 int nrhs, int64_t sstride, int64_t dstride) {
  const int d = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= len[d]) return;
  const int pi = perm[poff[d] + i];
  for (int o = 0; o < nrhs; o++) {
    if (scatter) dst[doff[d] + pi + o * dstride] = src[soff[d] + i + o * sstride];
    else dst[doff[d] + i + o * dstride] = src[soff[d] + pi + o * sstride];
  }
}

hipError_t launch_permute(hipStream_t st, int nsub, int maxlen, int scatter, const int* len, const int* perm,
                          const int64_t* poff, const double* src, const int64_t* soff, double* dst,
                          const int64_t* doff, int nrhs, int64_t sstride, int64_t dstride) {
  hipLaunchKernelGGL(k_permute, dim3((maxlen + 255) / 256, nsub), dim3(256), 0, st, scatter, len, perm, poff, src,
                     soff, dst, doff, nrhs, sstride, dstride);
  return hipGetLastError();
}

// rectangular copy dst[r + c*ldd] = src[r + c*lds], r < rows, c < cols
__global__ __launch_bounds__(256) void k_copy_cols(const int* __restrict__ rows, const int* __restrict__ cols,
                                                   const double* __restrict__ src, const int64_t* __restrict__ soff,
                                                   const int* __restrict__ lds, double* __restrict__ dst,
                                                   const int64_t* __restrict__ doff, const int* __restrict__ ldd,
                                                   double scale) {
  const int mi = blockIdx.y;
  const int c = blockIdx.x;
  if (c >= cols[mi]) return;
  const double* S = src + soff[mi] + (int64_t)c * lds[mi];
  double* D = dst + doff[mi] + (int64_t)c * ldd[mi];
  const int nr = rows[mi];
  for (int r = threadIdx.x; r < nr; r += 256) D[r] = scale * S[r];
}

hipError_t launch_copy_cols(hipStream_t st, int nmat, int maxcols, const int* rows, const int* cols,
                            const double* src, const int64_t* soff, const int* lds, double* dst,
                            const int64_t* doff, const int* ldd, double scale) {
  hipLaunchKernelGGL(k_copy_cols, dim3(maxcols, nmat), dim3(256), 0, st, rows, cols, src, soff, lds, dst, doff,
                     ldd, scale);
  return hipGetLastError();
}

// M[r + c*ld] *= s for row0 <= c <= r < n
__global__ __launch_bounds__(256) void k_scale_lower(const int* __restrict__ row0, const int* __restrict__ nn,
                                                     double* __restrict__ M, const int64_t* __restrict__ moff,
                                                     const int* __restrict__ ld, double s) {
  const int mi = blockIdx.y;
  const int c = row0[mi] + blockIdx.x;
  const int n = nn[mi];
  if (c >= n) return;
  double* col = M + moff[mi] + (int64_t)c * ld[mi];
  for (int r = c + threadIdx.x; r < n; r += 256) col[r] *= s;
}

hipError_t launch_scale_lower(hipStream_t st, int nmat, int maxcols, const int* row0, const int* n, double* M,
                              const int64_t* moff, const int* ld, double s) {
  hipLaunchKernelGGL(k_scale_lower, dim3(maxcols, nmat), dim3(256), 0, st, row0, n, M, moff, ld, s);
  return hipGetLastError();
}

// Y[p + k*ldy] = sum_q A[p + q*lda] X[q + k*ldx], p < n, k < ncol (8 per z slice);
// the X entries are wave-uniform loads (cached), A is streamed down its columns.
__global__ __launch_bounds__(256) void k_gemm_nn(const int* __restrict__ nn, const int* __restrict__ nin,
                                                 const int* __restrict__ ncol,
                                                 const double* __restrict__ A, const int64_t* __restrict__ aoff,
                                                 const int* __restrict__ lda, const double* __restrict__ X,
                                                 const int64_t* __restrict__ xoff, const int* __restrict__ ldx,
                                                 double* __restrict__ Y, const int64_t* __restrict__ yoff,
                                                 const int* __restrict__ ldy, double alpha) {
  const int mi = blockIdx.z;
  const int n = nn[mi], nc = ncol[mi], ni = nin ? nin[mi] : nn[mi];
  const int k0 = blockIdx.y * 8;
  if (k0 >= nc) return;
  const int kc = (nc - k0) < 8 ? (nc - k0) : 8;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double* Xm = X + xoff[mi];
  const int64_t lx = ldx[mi];
  const double* Am = A + aoff[mi];
  const int64_t la = lda[mi];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = 0; q < ni; q++) {
    const double a = Am[p + (int64_t)q * la];
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k < kc) acc[k] += a * Xm[q + (int64_t)(k0 + k) * lx];
  }
  double* Ym = Y + yoff[mi];
  const int64_t ly = ldy[mi];
  for (int k = 0; k < kc; k++) Ym[p + (int64_t)(k0 + k) * ly] = alpha * acc[k];
}

hipError_t launch_gemm_nn(hipStream_t st, int nmat, int maxn, int maxcol, const int* n, const int* nin,
                          const int* ncol, const double* A, const int64_t* aoff, const int* lda, const double* X,
                          const int64_t* xoff, const int* ldx, double* Y, const int64_t* yoff, const int* ldy,
                          double alpha) {
  if (maxcol < 1) return hipSuccess;
  hipLaunchKernelGGL(k_gemm_nn, dim3((maxn + 255) / 256, (maxcol + 7) / 8, nmat), dim3(256), 0, st, n, nin, ncol, A,
                     aoff, lda, X, xoff, ldx, Y, yoff, ldy, alpha);
  return hipGetLastError();
}

// y_i = A_i x1_i + beta * B_i x2_i with A_i (m x m1), B_i (m x m2) column major,
// ld even; x1 then x2 staged through one LDS buffer; each thread owns two
// consecutive rows.  Used for x_r = K_rr^-1 v_r - (K_rr^-1 K_rb) x_b.
__global__ __launch_bounds__(256) void k_gemv_two(const int* __restrict__ mm, const int* __restrict__ mm1,
                                                  const int* __restrict__ mm2, const double* __restrict__ A,
                                                  const int64_t* __restrict__ aoff, const int* __restrict__ lda,
                                                  const double* __restrict__ B, const int64_t* __restrict__ boff,
                                                  const int* __restrict__ ldb, const double* __restrict__ x1,
                                                  const int64_t* __restrict__ x1off, const double* __restrict__ x2,
                                                  const int64_t* __restrict__ x2off, double beta,
                                                  double* __restrict__ y, const int64_t* __restrict__ yoff) {
  extern __shared__ double xs[];
  const int d = blockIdx.y;
  const int m = mm[d], m1 = mm1[d], m2 = mm2[d];
  const int r0 = blockIdx.x * 512;
  if (r0 >= m) return;
  const int r = r0 + 2 * threadIdx.x;
  const bool two = r + 1 < m, one = r < m;
  double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
  for (int p = threadIdx.x; p < m1; p += 256) xs[p] = x1[x1off[d] + p];
  __syncthreads();
  if (one) {
    const double* Ad = A + aoff[d] + r;
    const int64_t la = lda[d];
    if (two) {
      for (int j = 0; j < m1; j++) {
        const double2 c = *reinterpret_cast<const double2*>(Ad + (int64_t)j * la);
        a0 += c.x * xs[j];
        a1 += c.y * xs[j];
      }
    } else {
      for (int j = 0; j < m1; j++) a0 += Ad[(int64_t)j * la] * xs[j];
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < m2; p += 256) xs[p] = x2[x2off[d] + p];
  __syncthreads();
  if (!one) return;
  const double* Bd = B + boff[d] + r;
  const int64_t lb = ldb[d];
  if (two) {
    for (int j = 0; j < m2; j++) {
      const double2 c = *reinterpret_cast<const double2*>(Bd + (int64_t)j * lb);
      b0 += c.x * xs[j];
      b1 += c.y * xs[j];
    }
    y[yoff[d] + r] = a0 + beta * b0;
    y[yoff[d] + r + 1] = a1 + beta * b1;
  } else {
    for (int j = 0; j < m2; j++) b0 += Bd[(int64_t)j * lb] * xs[j];
    y[yoff[d] + r] = a0 + beta * b0;
  }
}

hipError_t launch_gemv_two(hipStream_t st, int nsub, int maxm, int maxlds, const int* m, const int* m1,
                           const int* m2, const double* A, const int64_t* aoff, const int* lda, const double* B,
                           const int64_t* boff, const int* ldb, const double* x1, const int64_t* x1off,
                           const double* x2, const int64_t* x2off, double beta, double* y, const int64_t* yoff) {
  if ((size_t)maxlds * sizeof(double) > 65536) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gemv_two, dim3((maxm + 511) / 512, nsub), dim3(256), (size_t)maxlds * sizeof(double), st, m, m1,
                     m2, A, aoff, lda, B, boff, ldb, x1, x1off, x2, x2off, beta, y, yoff);
  return hipGetLastError();
}

// out[i + j*ldo] = alpha * sum_p X[p + i*ldx] Y[p + j*ldy], i < ni, j < nj  (one wave per pair)
__global__ __launch_bounds__(256) void k_tdots(const int* __restrict__ np, const int* __restrict__ ni,
                                               const int* __restrict__ nj, const double* __restrict__ X,
                                               const int64_t* __restrict__ xoff, const int* __restrict__ ldx,
                                               const double* __restrict__ Y, const int64_t* __restrict__ yoff,
                                               const int* __restrict__ ldy, double* __restrict__ out,
                                               const int64_t* __restrict__ ooff, const int* __restrict__ ldo,
                                               double alpha) {
  const int mi = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int a = ni[mi], b = nj[mi];
  if (pair >= a * b) return;
  const int i = pair % a, j = pair / a;
  const double* x = X + xoff[mi] + (int64_t)i * ldx[mi];
  const double* y = Y + yoff[mi] + (int64_t)j * ldy[mi];
  const int n = np[mi];
  double s = 0.0;
  for (int p = lane; p < n; p += 64) s += x[p] * y[p];
  s = wsum(s);
  if (lane == 0) out[ooff[mi] + i + (int64_t)j * ldo[mi]] = alpha * s;
}

hipError_t launch_tdots(hipStream_t st, int nmat, int maxpairs, const int* np, const int* ni, const int* nj,
                        const double* X, const int64_t* xoff, const int* ldx, const double* Y, const int64_t* yoff,
                        const int* ldy, double* out, const int64_t* ooff, const int* ldo, double alpha) {
  if (maxpairs < 1) return hipSuccess;
  hipLaunchKernelGGL(k_tdots, dim3((maxpairs + 3) / 4, nmat), dim3(256), 0, st, np, ni, nj, X, xoff, ldx, Y, yoff,
                     ldy, out, ooff, ldo, alpha);
  return hipGetLastError();
}

// v_i[border[p]] -= (B_i^T lam)[p]   and   v_i[border[p]] -= (Bc_i^T lam_c)[p]
__global__ __launch_bounds__(256) void k_sub_bt(const int* __restrict__ nr, const int* __restrict__ ns,
                                                const int* __restrict__ perm, const int64_t* __restrict__ noff,
                                                const int64_t* __restrict__ boff, const int* __restrict__ gptr,
                                                const int* __restrict__ gidx, const double* __restrict__ lam,
                                                const double* __restrict__ BcT, const int64_t* __restrict__ coff,
                                                const int* __restrict__ ldc, const int* __restrict__ nlc,
                                                const int64_t* __restrict__ lcoff, const int* __restrict__ lcmap,
                                                const double* __restrict__ la, double* __restrict__ v) {
  const int d = blockIdx.y;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int n = ns[d];
  if (p >= n) return;
  double s = 0.0;
  if (lam) {
    const int64_t q = boff[d] + p;
    for (int e = gptr[q]; e < gptr[q + 1]; e++) {
      const int code = gidx[e];
      const int r = code >= 0 ? code : ~code;
      s += code >= 0 ? lam[r] : -lam[r];
    }
  }
  if (BcT) {
    const double* B = BcT + coff[d];
    const int64_t l = ldc[d];
    const int nc = nlc[d];
    for (int k = 0; k < nc; k++) s += B[p + (int64_t)k * l] * la[lcmap[lcoff[d] + k]];
  }
  v[noff[d] + perm[noff[d] + nr[d] + p]] -= s;
}

hipError_t launch_sub_bt(hipStream_t st, int nsub, int maxns, const int* nr, const int* ns, const int* perm,
                         const int64_t* noff, const int64_t* boff, const int* gptr, const int* gidx,
                         const double* lam, const double* BcT, const int64_t* coff, const int* ldc, const int* nlc,
                         const int64_t* lcoff, const int* lcmap, const double* la, double* v) {
  hipLaunchKernelGGL(k_sub_bt, dim3((maxns + 255) / 256, nsub), dim3(256), 0, st, nr, ns, perm, noff, boff, gptr,
                     gidx, lam, BcT, coff, ldc, nlc, lcoff, lcmap, la, v);
  return hipGetLastError();
}

// x_i += R_i (la[alpha_idx_i + k] + h_i[k]);  then u[gdof] = x where owned
__global__ __launch_bounds__(256) void k_recover(const int* __restrict__ nn, const int64_t* __restrict__ noff,
                                                 const double* __restrict__ R, const int64_t* __restrict__ roff,
                                                 const int64_t* __restrict__ alpha_idx,
                                                 const double* __restrict__ la, const double* __restrict__ h,
                                                 const double* __restrict__ x, const int* __restrict__ gdof,
                                                 const unsigned char* __restrict__ owner, double* __restrict__ u) {
  const int d = blockIdx.y;
  const int l = blockIdx.x * 256 + threadIdx.x;
  const int n = nn[d];
  if (l >= n) return;
  const double* Rd = R + roff[d];
  double s = x[noff[d] + l];
  for (int k = 0; k < 6; k++) s += Rd[l + (int64_t)k * n] * (la[alpha_idx[d] + k] + h[6 * d + k]);
  if (owner[noff[d] + l]) u[gdof[noff[d] + l]] = s;
}

hipError_t launch_recover(hipStream_t st, int nsub, int maxn, const int* n, const int64_t* noff, const double* R,
                          const int64_t* roff, const int64_t* alpha_idx, const double* la, const double* h,
                          const double* x, const int* gdof, const unsigned char* owner, double* u) {
  hipLaunchKernelGGL(k_recover, dim3((maxn + 255) / 256, nsub), dim3(256), 0, st, n, noff, R, roff, alpha_idx, la, h,
                     x, gdof, owner, u);
  return hipGetLastError();
}

// out[0] = sum_i x[i]^2 over n entries: one workgroup, each thread a fixed strided
// range, then a fixed-shape tree (deterministic)
__global__ __launch_bounds__(1024) void k_sqnorm(int64_t n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) s += x[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

hipError_t launch_sqnorm(hipStream_t st, int64_t n, const double* x, double* out) {
  hipLaunchKernelGGL(k_sqnorm, dim3(1), dim3(1024), 0, st, n, x, out);
  return hipGetLastError();
}

// Lumped nodal forces of a smoothed face traction t (c, kk, jj) on the x = 1 face
// into one packed local load vector f (zeroed before): for every subdomain on that
// face, f = t * (cnt * hy * hz * 0.25) at its face nodes, cnt the subdomain's face
// elements at the node -- the lumping of traction_loads (problem.cpp), the same
// IEEE operations in the same order.  meta per face subdomain: e0 y, ne y, e0 z,
// ne z, local nodes along x, local nodes in an x-y layer.
__global__ __launch_bounds__(256) void k_traction_lump(const int* __restrict__ meta, const int64_t* __restrict__ foff,
                                                       int ny1, int nz1, double hy, double hz,
                                                       const double* __restrict__ t, double* __restrict__ f) {
  const int q = blockIdx.x;
  const int* mq = meta + 6 * q;
  const int e0y = mq[0], ney = mq[1], e0z = mq[2], nez = mq[3], nn0 = mq[4], nn01 = mq[5];
  double* fd = f + foff[q];
  const int nodes = (ney + 1) * (nez + 1);
  for (int p = threadIdx.x; p < nodes; p += 256) {
    const int a = p % (ney + 1), b = p / (ney + 1);
    const int c1 = (a > 0) + (a < ney), c2 = (b > 0) + (b < nez);
    const double w = (double)(c1 * c2) * hy * hz * 0.25;
    const int ln = (nn0 - 1) + a * nn0 + b * nn01;
    for (int c = 0; c < 3; c++) fd[3 * ln + c] = t[((int64_t)c * nz1 + e0z + b) * ny1 + e0y + a] * w;
  }
}

hipError_t launch_traction_lump(hipStream_t st, int nface, const int* meta, const int64_t* foff, int ny1, int nz1,
                                double hy, double hz, const double* t, double* f) {
  if (nface <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_traction_lump, dim3(nface), dim3(256), 0, st, meta, foff, ny1, nz1, hy, hz, t, f);
  return hipGetLastError();
}

}  // namespace hfls
