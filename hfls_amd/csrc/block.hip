// Dual-space and small kernels of the block PCPG: k load cases advanced together
// (Cluster.cpp:1363-1458 per column), each with its own scalars and stopping
// test.  Layouts:
//   dual block    v[r*k + j]               (row r of the dual space, column j < k)
//   b-block       X[(obrow[d] + p)*kp + j]  rows padded to 64 per subdomain
//   cluster block la[(gcrow[cl] + l)*kp + j]
//   local block   V[j*npacked + noff[d] + l] (column major packed local vectors)
// A column whose status is not 0 is frozen: its scalars and dual iterates are not
// updated, so every column's arithmetic is the same whether it is solved alone or
// inside a block.
#include <algorithm>

#include "kernels.hpp"

namespace hfls {

// workgroups of the grid-stride elementwise kernels: enough to cover every CU
// several times, few enough that a launch skipped through its gate costs ~us
static constexpr int64_t kStrideGrid = 4096;

static __device__ __forceinline__ double wsum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Signed row sums over the Bf^T lists of NQ consecutive b positions q0..q0+nq-1
// (entries gptr[q]..gptr[q+1]-1, code >= 0 for +1 at dual row code, ~code for -1):
// s[u] = sum_e sign_e f(r_e), added in entry order.  q0, nq and the entry indices
// are wave-uniform (scalar loads); the entries of the group are walked as one flat
// range, EB per step with every value load issued before any is used, so a wave
// keeps EB loads in flight instead of one per row.  gidx carries EB - 1 readable
// entries past its end (every r it holds is a dual row, so f(r) stays in bounds).
template <int NQ, int EB, class F>
__device__ __forceinline__ void bt_row_sums(int64_t q0, int nq, const int* __restrict__ gptr,
                                            const int* __restrict__ gidx, F f, double* s) {
  int eb[NQ + 1];
#pragma unroll
  for (int u = 0; u <= NQ; u++) eb[u] = gptr[q0 + (u < nq ? u : nq)];
#pragma unroll
  for (int u = 0; u < NQ; u++) s[u] = 0.0;
  const int e1 = eb[NQ];
  for (int e = eb[0]; e < e1; e += EB) {
    double val[EB];
#pragma unroll
    for (int i = 0; i < EB; i++) {
      const int code = gidx[e + i];  // past e1: a padding or next-group entry, loaded and dropped
      const double v = f(code >= 0 ? code : ~code);
      val[i] = code >= 0 ? v : -v;
    }
#pragma unroll
    for (int i = 0; i < EB; i++) {
      const int ei = e + i;
#pragma unroll
      for (int u = 0; u < NQ; u++)
        if (ei >= eb[u] && ei < eb[u + 1]) s[u] += val[i];
    }
  }
}

// xb[qrow[q]*kp + j] = sum_e sign_e * wgt[r_e] * v[r_e*k + j] (qrow[q] = obrow[d] + p for
// b position p of subdomain d); optional FP32 copy.
// A wave takes GQ consecutive b positions per step (lanes = columns); the grid is
// capped so a gated launch is cheap.
constexpr int GQ = 8;
template <bool W>
__global__ __launch_bounds__(256) void kb_gather(int64_t nb, int k, int kp, const int* __restrict__ qrow,                                                  const int* __restrict__ gptr, const int* __restrict__ gidx,
                                                 const double* __restrict__ v, const double* __restrict__ wgt,
                                                 double* __restrict__ xb, float* __restrict__ xb32,
                                                 const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int j = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  const int64_t ngroups = (nb + GQ - 1) / GQ;
  if (j >= k) return;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: index loads go scalar
  for (int64_t grp = (int64_t)blockIdx.x * 4 + wv; grp < ngroups; grp += nwaves) {
    const int64_t q0 = grp * GQ;
    const int nq = (int)std::min<int64_t>(GQ, nb - q0);
    double s[GQ];
    bt_row_sums<GQ, 8>(q0, nq, gptr, gidx,
                       [&](int r) {
                         const double val = v[(int64_t)r * k + j];
                         return W ? val * wgt[r] : val;
                       },
                       s);
#pragma unroll
    for (int u = 0; u < GQ; u++) {
      if (u >= nq) break;
      const int64_t o = (int64_t)qrow[q0 + u] * kp + j;
      if (xb) xb[o] = s[u];
      if (xb32) xb32[o] = (float)s[u];
    }
  }
}

hipError_t launch_kb_gather(hipStream_t st, int64_t nb, int k, int kp, const int* qrow, const int* gptr, const int* gidx, const double* v,
                            const double* wgt, double* xb, float* xb32, const int* gate) {
  if (k > 64) return hipErrorInvalidValue;
  const int64_t groups = (nb + GQ - 1) / GQ;
  const dim3 grid((unsigned)std::min<int64_t>((groups + 3) / 4, kStrideGrid));
  if (wgt)
    hipLaunchKernelGGL(kb_gather<true>, grid, dim3(256), 0, st, nb, k, kp, qrow, gptr, gidx, v, wgt, xb,
                       xb32, gate);
  else
    hipLaunchKernelGGL(kb_gather<false>, grid, dim3(256), 0, st, nb, k, kp, qrow, gptr, gidx, v, wgt, xb,
                       xb32, gate);
  return hipGetLastError();
}

// G^T x of the projector without an interface block in HBM, in chunks of at most
// GTC b positions of one subdomain (chunk c: subdomain csub[c], positions
// cp0[c] .. min(cp0[c] + GTC, n_s)), so large and small subdomains spread evenly over
// the device: part[(6c + m)*kp + j] = sum over the chunk's Bf^T entries e of
// R_b[p_e][m] * sign_e * x[r_e][j] (R_b column major at Rb[offRb[d] + p + m*n_s], p_e =
// gpos[e] the entry's b position).  The chunk's R_b rows, entry codes and positions
// are first staged in LDS with coalesced loads; its entries are then split into 4
// contiguous ranges, one per wave, walked GXB at a time with every x load of a batch
// issued before any is used (codes, positions and R_b rows are LDS broadcasts).  The
// 4 wave sums are added in a fixed order; kb_gtx_sum then adds a subdomain's chunks
// in ascending order: out[(6d + m)*kp + j].  A column's value does not depend on the
// other columns.
constexpr int GTC = 256;
constexpr int GXB = 16;
constexpr int GTE = 8 * GTC;  // entries of a chunk: at most 8 dual rows meet at a b position
__global__ __launch_bounds__(256) void kb_gtx(int k, int kp, const int* __restrict__ ns,
                                              const int* __restrict__ csub, const int* __restrict__ cp0,
                                              const int64_t* __restrict__ boff, const int* __restrict__ gptr,
                                              const int* __restrict__ gidx, const int* __restrict__ gpos,
                                              const double* __restrict__ Rb, const int64_t* __restrict__ offRb,
                                              const double* __restrict__ x, double* __restrict__ part,
                                              const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double sR[GTC * 6];  // R_b row p - pbeg at 6 (p - pbeg); reused for the wave sums
  __shared__ int scode[GTE];
  __shared__ unsigned char spos[GTE];
  const int c = blockIdx.x, j = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int d = csub[c], n = ns[d];
  const int pbeg = cp0[c], pend = std::min(n, pbeg + GTC), np = pend - pbeg;
  const double* R = Rb + offRb[d];
  const int e0 = gptr[boff[d] + pbeg], e1 = gptr[boff[d] + pend];
  const int ne = std::min(e1 - e0, GTE);
  for (int t = threadIdx.x; t < 6 * np; t += 256) {
    const int m = t / np, p = t - m * np;
    sR[6 * p + m] = R[pbeg + p + (int64_t)m * n];
  }
  for (int t = threadIdx.x; t < ne; t += 256) {
    scode[t] = gidx[e0 + t];
    spos[t] = (unsigned char)(gpos[e0 + t] - pbeg);
  }
  __syncthreads();
  const int ea = (int)((int64_t)ne * w / 4), eb = (int)((int64_t)ne * (w + 1) / 4);
  const int jc = j < k ? j : 0;  // lanes past k read column 0 and drop it
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int e = ea; e < eb; e += GXB) {
    double xv[GXB];
#pragma unroll
    for (int i = 0; i < GXB; i++) {
      const int code = scode[std::min(e + i, eb - 1)];  // past eb: a valid entry, loaded and dropped
      const double v = x[(int64_t)(code >= 0 ? code : ~code) * k + jc];
      xv[i] = code >= 0 ? v : -v;
    }
#pragma unroll
    for (int i = 0; i < GXB; i++) {
      if (e + i < eb) {
        const double* r6 = sR + 6 * spos[e + i];
#pragma unroll
        for (int m = 0; m < 6; m++) acc[m] += r6[m] * xv[i];
      }
    }
  }
  __syncthreads();
  double* red = sR;  // [w][m][64]
#pragma unroll
  for (int m = 0; m < 6; m++) red[(w * 6 + m) * 64 + j] = acc[m];
  __syncthreads();
  for (int t = threadIdx.x; t < 6 * 64; t += 256) {
    const int m = t >> 6, jj = t & 63;
    if (jj >= k) continue;
    part[((int64_t)6 * c + m) * kp + jj] =
        red[(0 * 6 + m) * 64 + jj] + red[(1 * 6 + m) * 64 + jj] + red[(2 * 6 + m) * 64 + jj] +
        red[(3 * 6 + m) * 64 + jj];
  }
}

__global__ __launch_bounds__(384) void kb_gtx_sum(int k, int kp, const int* __restrict__ cbeg,
                                                  const double* __restrict__ part, double* __restrict__ out,
                                                  const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int d = blockIdx.x, m = threadIdx.x >> 6, j = threadIdx.x & 63;
  if (j >= k) return;
  double s = 0.0;
  for (int c = cbeg[d]; c < cbeg[d + 1]; c++) s += part[((int64_t)6 * c + m) * kp + j];
  out[((int64_t)6 * d + m) * kp + j] = s;
}

hipError_t launch_kb_gtx(hipStream_t st, int nsub, int nchunk, int k, int kp, const int* ns, const int* csub,
                         const int* cp0, const int* cbeg, const int64_t* boff, const int* gptr, const int* gidx,
                         const int* gpos, const double* Rb, const int64_t* offRb, const double* x, double* part,
                         double* out, const int* gate) {
  if (k > 64 || k > kp || nchunk < 1 || !gpos) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kb_gtx, dim3(nchunk), dim3(256), 0, st, k, kp, ns, csub, cp0, boff, gptr, gidx, gpos, Rb, offRb,
                     x, part, gate);
  hipLaunchKernelGGL(kb_gtx_sum, dim3(nsub), dim3(384), 0, st, k, kp, cbeg, part, out, gate);
  return hipGetLastError();
}

// plain copy of n doubles (16-byte accesses), the reference point for HBM traffic
// counters: it reads and writes exactly 8 n bytes each
__global__ __launch_bounds__(256) void kb_copy(int64_t n2, const double2* __restrict__ a, double2* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n2) b[t] = a[t];
}

hipError_t launch_kb_copy(hipStream_t st, int64_t n, const double* a, double* b) {
  if (n % 2) return hipErrorInvalidValue;
  const int64_t n2 = n / 2;
  hipLaunchKernelGGL(kb_copy, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, n2, (const double2*)a,
                     (double2*)b);
  return hipGetLastError();
}

// elementwise FP64 -> FP32 copy (tile stores of the same layout)
__global__ __launch_bounds__(256) void kb_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

hipError_t launch_kb_to_f32(hipStream_t st, int64_t n, const double* a, float* b) {
  hipLaunchKernelGGL(kb_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b);
  return hipGetLastError();
}

// a[(64d + m)*kp + j] = sum_l R[l + m*n] * V[j*npacked + noff_d + l]   (R_i^T v_i, full local)
// one workgroup per (subdomain, column): v is read once for the six products
__global__ __launch_bounds__(256) void kb_rtf(int k, int kp, int64_t npacked, const int* __restrict__ nn,
                                              const int64_t* __restrict__ noff, const double* __restrict__ R,
                                              const int64_t* __restrict__ roff, const double* __restrict__ V,
                                              double* __restrict__ a) {
  const int d = blockIdx.x;
  const int j = blockIdx.y;
  const int n = nn[d];
  __shared__ double sh[6][4];
  const double* Rd = R + roff[d];
  const double* v = V + (int64_t)j * npacked + noff[d];
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int l = threadIdx.x; l < n; l += 256) {
    const double x = v[l];
#pragma unroll
    for (int m = 0; m < 6; m++) s[m] += Rd[l + (int64_t)m * n] * x;
  }
#pragma unroll
  for (int m = 0; m < 6; m++) {
    const double t = wsum64(s[m]);
    if ((threadIdx.x & 63) == 0) sh[m][threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x < 6)
    a[((int64_t)64 * d + threadIdx.x) * kp + j] =
        sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
}

hipError_t launch_kb_rtf(hipStream_t st, int nsub, int k, int kp, int64_t npacked, const int* n, const int64_t* noff,
                         const double* R, const int64_t* roff, const double* V, double* a) {
  hipLaunchKernelGGL(kb_rtf, dim3(nsub, k), dim3(256), 0, st, k, kp, npacked, n, noff, R, roff, V, a);
  return hipGetLastError();
}

// gc[(gcrow[cl] + i)*kp + j]:  i < nl: c rows (cluster constraint sums), else -a rows
//   c rows: cr[lcpos[src]*kp + j] (+ second contribution), a rows: a[(abase[d] + m)*kp + j]
__global__ __launch_bounds__(256) void kb_cluster_gc(int k, int kp, const int* __restrict__ gc_src,
                                                     const int64_t* __restrict__ gc_off, const int* __restrict__ nlcl,
                                                     const int* __restrict__ nrcl, const int* __restrict__ d0,
                                                     const double* __restrict__ cr, const int64_t* __restrict__ lcpos,
                                                     const double* __restrict__ a, const int64_t* __restrict__ abase,
                                                     double* __restrict__ gc, const int64_t* __restrict__ gcrow, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int cl = blockIdx.y;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int i = (int)(t / k), j = (int)(t % k);
  const int nl = nlcl[cl], nr = nrcl[cl];
  if (i >= nl + nr) return;
  double v;
  if (i < nl) {
    const int* s = gc_src + 2 * (gc_off[cl] + i);
    v = cr[lcpos[s[0]] * kp + j];
    if (s[1] >= 0) v += cr[lcpos[s[1]] * kp + j];
  } else {
    const int m = i - nl;
    const int d = d0[cl] + m / 6;
    v = -a[(abase[d] + m % 6) * kp + j];
  }
  gc[(gcrow[cl] + i) * kp + j] = v;
}

hipError_t launch_kb_cluster_gc(hipStream_t st, int ncl, int maxlen, int k, int kp, const int* gc_src,
                                const int64_t* gc_off, const int* nlcl, const int* nrcl, const int* d0,
                                const double* cr, const int64_t* lcpos, const double* a, const int64_t* abase,
                                double* gc, const int64_t* gcrow, const int* gate) {
  const int64_t n = (int64_t)maxlen * k;
  hipLaunchKernelGGL(kb_cluster_gc, dim3((unsigned)((n + 255) / 256), ncl), dim3(256), 0, st, k, kp, gc_src, gc_off,
                     nlcl, nrcl, d0, cr, lcpos, a, abase, gc, gcrow, gate);
  return hipGetLastError();
}

// Coefficient block of the correction y_b += CRQ coef (a tile GEMM with the
// compact n_s x (nlc + 18) block CRQ = [C | R_b | Q | Yh], ld ldc):
//   coef[(crow[d] + m)*kp + j] = -la[lcrow_m][j]            m < nlc  (-C lam_c)
//                              = with_R * la[arow_d + m'][j]         (R_b alpha)
//                              = defl * (Yh^T x)[m'][j], defl * (Q^T x)[m'][j]
// where (Q^T x, Yh^T x) are rows crow[d] + nlc + 6.. and + 12.. of cr
__global__ __launch_bounds__(256) void kb_coef(int k, int kp, const int* __restrict__ nlc,
                                               const int64_t* __restrict__ lcoff, const int64_t* __restrict__ lcrow,
                                               const int64_t* __restrict__ arow, const double* __restrict__ la,
                                               int with_R, int defl, const double* __restrict__ cr,
                                               const int64_t* __restrict__ crow, double* __restrict__ coef,
                                               const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int d = blockIdx.y;
  const int nc = nlc[d];
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int m = (int)(t / k), j = (int)(t % k);
  if (m >= nc + 18) return;
  double v;
  if (m < nc) v = -la[lcrow[lcoff[d] + m] * kp + j];
  else if (m < nc + 6) v = with_R ? la[(arow[d] + (m - nc)) * kp + j] : 0.0;
  else if (m < nc + 12) v = defl ? cr[(crow[d] + m + 6) * kp + j] : 0.0;
  else v = defl ? cr[(crow[d] + m - 6) * kp + j] : 0.0;
  coef[(crow[d] + m) * kp + j] = v;
}

hipError_t launch_kb_coef(hipStream_t st, int nsub, int maxrows, int k, int kp, const int* nlc, const int64_t* lcoff,
                          const int64_t* lcrow, const int64_t* arow, const double* la, int with_R, int defl,
                          const double* cr, const int64_t* crow, double* coef, const int* gate) {
  const int64_t n = (int64_t)maxrows * k;
  hipLaunchKernelGGL(kb_coef, dim3((unsigned)((n + 255) / 256), nsub), dim3(256), 0, st, k, kp, nlc, lcoff, lcrow, arow,
                     la, with_R, defl, cr, crow, coef, gate);
  return hipGetLastError();
}

// X[(obrow[d] + p)*kp + m] = A[p + m*lda] (dir 0) or the reverse (dir 1), p < n_s, m < ncol
__global__ __launch_bounds__(256) void kb_cols(int dir, int ncol, int kp, const int* __restrict__ ns,
                                               const int64_t* __restrict__ obrow, double* __restrict__ A,
                                               const int64_t* __restrict__ aoff, const int* __restrict__ lda,
                                               double* __restrict__ X) {
  const int d = blockIdx.y;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int p = (int)(t / ncol), m = (int)(t % ncol);
  if (p >= ns[d]) return;
  double* a = A + aoff[d] + p + (int64_t)m * lda[d];
  double* x = X + (obrow[d] + p) * kp + m;
  if (dir == 0) *x = *a;
  else *a = *x;
}

hipError_t launch_kb_cols(hipStream_t st, int nsub, int maxns, int dir, int ncol, int kp, const int* ns,
                          const int64_t* obrow, double* A, const int64_t* aoff, const int* lda, double* X) {
  const int64_t n = (int64_t)maxns * ncol;
  hipLaunchKernelGGL(kb_cols, dim3((unsigned)((n + 255) / 256), nsub), dim3(256), 0, st, dir, ncol, kp, ns, obrow, A,
                     aoff, lda, X);
  return hipGetLastError();
}

// FP32 tiles of the deflated symmetric block A - Q Yh^T - Yh Q^T (packed lower
// tiles as the FP64 source); Q, Yh: columns nlc+6.., nlc+12.. of CRQ
__global__ __launch_bounds__(256) void kb_deflate_f32(const int* __restrict__ ns, const double* __restrict__ P,
                                                      const int64_t* __restrict__ poff, const double* __restrict__ CRQ,
                                                      const int64_t* __restrict__ coff, const int* __restrict__ ldc,
                                                      const int* __restrict__ nlc, float* __restrict__ P32) {
  const int d = blockIdx.y;
  const int n = ns[d];
  const int T = (n + 63) / 64;
  const int q = blockIdx.x;
  if (q >= T * (T + 1) / 2) return;
  int tj = 0, rem = q;
  while (rem >= T - tj) {
    rem -= T - tj;
    tj++;
  }
  const int ti = tj + rem;
  const int64_t idx = ((int64_t)tj * T - (int64_t)tj * (tj - 1) / 2 + (ti - tj)) * 4096;
  const double* src = P + poff[d] + idx;
  float* dst = P32 + poff[d] + idx;
  const double* Q = CRQ + coff[d] + (int64_t)(nlc[d] + 6) * ldc[d];
  const double* Y = CRQ + coff[d] + (int64_t)(nlc[d] + 12) * ldc[d];
  const int64_t l = ldc[d];
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = 64 * ti + (e & 63), c = 64 * tj + (e >> 6);
    double v = 0.0;
    if (r < n && c < n) {
      v = src[e];
      for (int m = 0; m < 6; m++) v -= Q[r + m * l] * Y[c + m * l] + Y[r + m * l] * Q[c + m * l];
    }
    dst[e] = (float)v;
  }
}

hipError_t launch_kb_deflate_f32(hipStream_t st, int nsub, int maxtiles, const int* ns, const double* P,
                                 const int64_t* poff, const double* CRQ, const int64_t* coff, const int* ldc,
                                 const int* nlc, float* P32) {
  hipLaunchKernelGGL(kb_deflate_f32, dim3(maxtiles, nsub), dim3(256), 0, st, ns, P, poff, CRQ, coff, ldc, nlc, P32);
  return hipGetLastError();
}

// out[r*k+j] = wgt[r]*(y_a[pos_a][j] - y_b[pos_b][j]) - sub[r*k+j]; per-column
// partial dots with dotw, 64 dual rows per workgroup: part[blk*k + j].  srow[2r] and
// srow[2r+1] are the b-block rows of dual row r's + and - entries (-1: no - entry).
// Wave g of a workgroup takes the 16 consecutive dual rows 64 blockIdx + 16 g + q in
// two batches of 8: the batch's 16 row indices are one scalar load, then every value
// load of the batch is issued before any is used.  TY: the b-block's element type
// (FP64, or the FP32 result of the mixed-precision GEMM).
template <typename TY>
__global__ __launch_bounds__(256) void kb_scatter(int64_t nlam, int k, int kp, const int* __restrict__ srow,
                                                  const TY* __restrict__ y, const double* __restrict__ wgt,
                                                  const double* __restrict__ sub, double* __restrict__ out,
                                                  const double* __restrict__ dotw, double* __restrict__ part,
                                                  const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double red[4 * 64];
  const int j = threadIdx.x & 63, g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double s = 0.0;
  if (j < k) {
    constexpr int SB = 8;
    const int64_t r0 = (int64_t)blockIdx.x * 64 + g * 16;
#pragma unroll
    for (int q0 = 0; q0 < 16; q0 += SB) {
      int sa[SB], sb[SB];
#pragma unroll
      for (int u = 0; u < SB; u++) {
        const int64_t r = r0 + q0 + u;
        sa[u] = r < nlam ? srow[2 * r] : -1;
        sb[u] = r < nlam ? srow[2 * r + 1] : -1;
      }
      double va[SB], vb[SB], ws[SB], sv[SB], dv[SB];
#pragma unroll
      for (int u = 0; u < SB; u++) {
        const int64_t r = r0 + q0 + u;
        va[u] = vb[u] = sv[u] = dv[u] = 0.0;
        ws[u] = 1.0;
        if (sa[u] >= 0) {
          va[u] = (double)y[(int64_t)sa[u] * kp + j];
          if (sb[u] >= 0) vb[u] = (double)y[(int64_t)sb[u] * kp + j];
          if (wgt) ws[u] = wgt[r];
          if (sub) sv[u] = sub[r * k + j];
          if (dotw) dv[u] = dotw[r * k + j];
        }
      }
#pragma unroll
      for (int u = 0; u < SB; u++) {
        const int64_t r = r0 + q0 + u;
        if (r >= nlam) break;
        double v = va[u];
        if (sb[u] >= 0) v -= vb[u];
        if (wgt) v *= ws[u];
        if (sub) v -= sv[u];
        out[r * k + j] = v;
        if (dotw) s += v * dv[u];
      }
    }
  }
  if (dotw) {
    red[g * 64 + j] = s;
    __syncthreads();
    if (g == 0 && j < k) part[(int64_t)blockIdx.x * k + j] = red[j] + red[64 + j] + red[128 + j] + red[192 + j];
  }
}

hipError_t launch_kb_scatter(hipStream_t st, int64_t nlam, int k, int kp, const int* srow, const double* y,
                             const double* wgt, const double* sub, double* out, const double* dotw, double* part,
                             const int* gate, const float* yf) {
  if (k > 64 || !srow) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((nlam + 63) / 64));
  if (yf)
    hipLaunchKernelGGL(kb_scatter<float>, grid, dim3(256), 0, st, nlam, k, kp, srow, yf, wgt, sub, out, dotw, part,
                       gate);
  else
    hipLaunchKernelGGL(kb_scatter<double>, grid, dim3(256), 0, st, nlam, k, kp, srow, y, wgt, sub, out, dotw, part,
                       gate);
  return hipGetLastError();
}

constexpr int GHR = 16;  // dual rows per wave
// out[r*k+j] = x[r*k+j] + sgn * (G_r,a . h_a[:,j] - G_r,b . h_b[:,j]); partial dots, where
// Gr[r*12 + m] = R_b,a[pos_a, m] and Gr[r*12 + 6 + m] = R_b,b[pos_b, m] (0 for Dirichlet
// rows) are the rows of G_f restricted to the two subdomains of dual row r.
// Workgroup g takes the 64 dual rows 64 g .. 64 g + 63, wave w the 16 rows 64 g + 16 w + q.
// The h rows of the (at most GH_SLOTS) subdomains of the group listed in
// gsub[g*GH_SLOTS ..] are first staged in LDS; rslot[r] = slot_a | slot_b << 8 names the
// LDS slots of row r's two subdomains (b = a for a row with one subdomain, whose
// half of G is zero), 255 a subdomain without a slot, read from h directly.  The x
// and dotv loads of a wave are issued first, so a row costs two vector loads, LDS
// reads and a store.
__global__ __launch_bounds__(256) void kb_gh(int64_t nlam, int k, const int* __restrict__ rsa,
                                             const int* __restrict__ rsb, const int* __restrict__ gsub,
                                             const int* __restrict__ rslot, const double* __restrict__ Gr,
                                             const double* __restrict__ h, const double* __restrict__ x,
                                             double* __restrict__ out, double sgn, const double* __restrict__ dotv,
                                             double* __restrict__ part, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double hs[GH_SLOTS * 6 * 64];
  __shared__ double red[4 * 64];
  const int j = threadIdx.x & 63, g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int t = threadIdx.x; t < GH_SLOTS * 6 * 64; t += 256) {
    const int sl = t / (6 * 64), mj = t - sl * 6 * 64, m = mj >> 6, jj = mj & 63;
    const int dsub = gsub[(int64_t)blockIdx.x * GH_SLOTS + sl];
    hs[t] = (dsub >= 0 && jj < k) ? h[((int64_t)6 * dsub + m) * k + jj] : 0.0;
  }
  __syncthreads();
  double s = 0.0;
  if (j < k) {
    const int64_t r0 = (int64_t)blockIdx.x * (4 * GHR) + g * GHR;
    const int nr = (int)std::max<int64_t>(0, std::min<int64_t>(GHR, nlam - r0));
    double xv[GHR], dv[GHR];
#pragma unroll
    for (int q = 0; q < GHR; q++) {
      xv[q] = (x && q < nr) ? x[(r0 + q) * k + j] : 0.0;
      dv[q] = (dotv && q < nr) ? dotv[(r0 + q) * k + j] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < GHR; q++) {
      if (q < nr) {
        const int64_t r = r0 + q;
        const int sl = rslot[r];
        const int sa = sl & 255, sb = sl >> 8;
        double ha[6], hb[6];
        if (sa != 255) {
#pragma unroll
          for (int m = 0; m < 6; m++) ha[m] = hs[(sa * 6 + m) * 64 + j];
        } else {
          const int a = rsa[r];
#pragma unroll
          for (int m = 0; m < 6; m++) ha[m] = h[((int64_t)6 * a + m) * k + j];
        }
        if (sb != 255) {
#pragma unroll
          for (int m = 0; m < 6; m++) hb[m] = hs[(sb * 6 + m) * 64 + j];
        } else {
          const int b0 = rsb[r];
          const int b = b0 >= 0 ? b0 : rsa[r];
#pragma unroll
          for (int m = 0; m < 6; m++) hb[m] = h[((int64_t)6 * b + m) * k + j];
        }
        const double* G = Gr + r * 12;
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < 6; m++) v += G[m] * ha[m];
#pragma unroll
        for (int m = 0; m < 6; m++) v -= G[6 + m] * hb[m];
        const double o = xv[q] + sgn * v;
        out[r * k + j] = o;
        s += o * dv[q];
      }
    }
  }
  if (dotv) {
    red[g * 64 + j] = s;
    __syncthreads();
    if (g == 0 && j < k) part[(int64_t)blockIdx.x * k + j] = red[j] + red[64 + j] + red[128 + j] + red[192 + j];
  }
}

hipError_t launch_kb_gh(hipStream_t st, int64_t nlam, int k, const int* rsa, const int* rsb, const int* gsub,
                        const int* rslot, const double* Gr, const double* h, const double* x, double* out, double sgn,
                        const double* dotv, double* part, const int* gate) {
  if (k > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kb_gh, dim3((unsigned)((nlam + 4 * GHR - 1) / (4 * GHR))), dim3(256), 0, st, nlam, k, rsa, rsb,
                     gsub, rslot, Gr, h, x, out, sgn, dotv, part, gate);
  return hipGetLastError();
}

// per column j: y = sum_{i in cl} H_i^T v_i ; alpha = invGtG y ; h_i = sgn * H_i alpha_cl(i)
// v_i rows at v[(vrow[i] + a)*kp + j] (vrow[i] = 6 i when vrow is null); one
// workgroup per column
__global__ __launch_bounds__(256) void kb_proj_coarse(int nsub, int ncol, int k, int kp, const double* __restrict__ v,
                                                      const int64_t* __restrict__ vrow, const double* __restrict__ H,
                                                      const int* __restrict__ cl, const double* __restrict__ invGtG,
                                                      double* __restrict__ h, double* __restrict__ alpha, double sgn,
                                                      const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int j = blockIdx.x;
  constexpr int CH = 256;  // subdomains per chunk of the first product
  __shared__ double y[1024];
  __shared__ double al[1024];
  __shared__ double z[CH * 6];
  const int ncl = ncol / 6, nsc = nsub / ncl;
  for (int c = threadIdx.x; c < ncol; c += 256) y[c] = 0.0;
  // y = H^T v per cluster, in chunks: z_d = H_d^T v_d for CH subdomains in
  // parallel, then each y entry adds its cluster's z_d of the chunk in order
  for (int d0 = 0; d0 < nsub; d0 += CH) {
    for (int t = threadIdx.x; t < CH * 6; t += 256) {
      const int d = d0 + t / 6, kk = t % 6;
      if (d >= nsub) break;
      const double* Hd = H + 36 * d;
      const int64_t row = vrow ? vrow[d] : (int64_t)6 * d;
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 6; a++) s += Hd[a * 6 + kk] * v[(row + a) * kp + j];
      z[t] = s;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncol; c += 256) {
      const int ccl = c / 6, kk = c % 6;
      const int da = max(d0, ccl * nsc), db = min(min(d0 + CH, nsub), (ccl + 1) * nsc);
      double s = 0.0;
      for (int d = da; d < db; d++) s += z[(d - d0) * 6 + kk];
      y[c] += s;
    }
    __syncthreads();
  }
  for (int c = threadIdx.x; c < ncol; c += 256) {
    double s = 0.0;
    for (int q = 0; q < ncol; q++) s += invGtG[c + (int64_t)q * ncol] * y[q];
    al[c] = s;
    if (alpha) alpha[(int64_t)c * k + j] = s;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 6 * nsub; q += 256) {
    const int d = q / 6, a = q % 6;
    const double* Hd = H + 36 * d + a * 6;
    const double* ad = al + 6 * cl[d];
    double s = 0.0;
    for (int b = 0; b < 6; b++) s += Hd[b] * ad[b];
    h[(int64_t)q * k + j] = sgn * s;
  }
}

hipError_t launch_kb_proj_coarse(hipStream_t st, int nsub, int ncol, int k, int kp, const double* v,
                                 const int64_t* vrow, const double* H, const int* cl, const double* invGtG, double* h,
                                 double* alpha, double sgn, const int* gate) {
  if (ncol > 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kb_proj_coarse, dim3(k), dim3(256), 0, st, nsub, ncol, k, kp, v, vrow, H, cl, invGtG, h, alpha,
                     sgn, gate);
  return hipGetLastError();
}

// Per-column reductions of the partial dots (fixed order) and the PCPG scalar
// logic (Cluster.cpp:1398-1448).  Column status: 0 running, 1 waiting for an FP64
// true-residual evaluation, 2 converged, 3 iteration limit, 4 empty slot,
// 5 joining (initialised by mode 2), 6 stopped on a non-finite residual (the
// column is frozen like a converged one; err records that some column hit it).
//   mode 0: wFw, rho = -gPz / wFw
//   mode 1: gamma, iteration count; the stopping test sqrt(gPz) < eps*norm0 gives
//           status 2 (status 1 when certification is on), a recursive residual
//           below rtol times the last true one gives status 1 (refinement point),
//           the iteration limit status 3; pending = some column has status 1
//   mode 2: initial gPz, norm0 of the joining columns
//   mode 3: (only while pending) for status-1 columns, from the FP64 true residual:
//           converged -> 2; limit -> 3; otherwise the iteration restarts from the
//           true residual (repl = 1: g <- g_true, w <- P M P g_true); pending is
//           cleared (the restart update that follows is gated by `restart`)
//   mode 4: record the true residual of every column (res)
// column sums of the partials: one workgroup per column, strided partial sums
// then a fixed-shape tree (the order depends on nparts only, not on other columns)
__global__ __launch_bounds__(256) void kb_colsum(int nparts, int k, const double* __restrict__ part, BScalars* sc,
                                                 const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int j = blockIdx.x;
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256) s += part[(int64_t)b * k + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sc->colsum[j] = red[0];
}

__global__ __launch_bounds__(64) void kb_reduce(int k, BScalars* sc, int mode, const int* __restrict__ gate) {
  const int j = threadIdx.x;
  __shared__ int any, run, rst, cert, nwait, nrun, maxw;
  if (gate && *gate == 0) return;              // uniform over the workgroup
  if (mode == 3 && sc->pending == 0) return;
  if (j == 0) any = run = rst = cert = nwait = nrun = maxw = 0;
  __syncthreads();
  if (j < k) {
    const double s = sc->colsum[j];
    const int stt = sc->done[j];
    if (mode == 0) {
      if (!stt) {
        sc->wFw[j] = s;
        sc->rho[j] = -sc->gPz[j] / s;
      }
    } else if (mode == 1) {
      sc->repl[j] = 0;
      if (!stt) {
        sc->gPz_prev[j] = sc->gPz[j];
        sc->gPz[j] = s;
        sc->gamma[j] = s / sc->gPz_prev[j];
        sc->it[j] += 1;
        const double r = sqrt(fabs(s));
        if (!isfinite(s)) {
          sc->done[j] = 6;
          sc->err = 1;
        } else if (!(r >= sc->eps * sc->norm0[j])) sc->done[j] = sc->cert ? 1 : 2;
        else if (r < sc->rtol * sc->ref[j]) sc->done[j] = 1;
        else if (sc->it[j] >= sc->max_iter) sc->done[j] = 3;
      }
      if (sc->done[j] == 1) {  // waiting (newly or still) for the FP64 evaluation
        sc->waitn[j] += 1;
        any = 1;
        atomicAdd(&nwait, 1);
        atomicMax(&maxw, sc->waitn[j]);
      } else if (sc->done[j] == 0) {
        atomicAdd(&nrun, 1);
      }
    } else if (mode == 2) {
      if (stt == 5) {
        sc->gPz[j] = s;
        sc->norm0[j] = sqrt(fabs(s));
        sc->ref[j] = sc->norm0[j];
        sc->it[j] = 0;
        sc->res[j] = 1.0;
        sc->repl[j] = 0;
        sc->ncert[j] = 0;
        sc->hgok[j] = 0;
        sc->waitn[j] = 0;
        if (!isfinite(s)) sc->err = 1;
        sc->done[j] = !isfinite(s) ? 6 : (!(s > 0.0) ? 2 : (sc->max_iter <= 0 ? 3 : 0));
      }
    } else if (mode == 3) {
      sc->repl[j] = 0;
      sc->certnow[j] = 0;
      if (stt == 1) {
        const double r = sqrt(fabs(s));
        sc->res[j] = sc->norm0[j] > 0 ? r / sc->norm0[j] : 0.0;
        sc->ncert[j] += 1;
        sc->waitn[j] = 0;
        if (!isfinite(s)) {
          sc->done[j] = 6;
          sc->err = 1;
        } else if (r < sc->eps * sc->norm0[j]) {
          sc->done[j] = 2;
          sc->certnow[j] = 1;
          sc->hgok[j] = 1;
          cert = 1;
        } else if (sc->it[j] >= sc->max_iter) {
          sc->done[j] = 3;
        } else {
          sc->repl[j] = 1;
          sc->gPz[j] = s;
          sc->ref[j] = r;
          sc->done[j] = 0;
          rst = 1;
        }
      }
    } else {
      if (stt != 4) sc->res[j] = sc->norm0[j] > 0 ? sqrt(fabs(s)) / sc->norm0[j] : 0.0;
    }
    if (sc->done[j] == 0 || sc->done[j] == 1) run = 1;
  }
  __syncthreads();
  if (j == 0) {
    // FP64 evaluations are batched: they run once a quarter of the live columns
    // wait, when no column is left iterating, or when a column has waited 3 sweeps
    // (a waiting column is frozen, so its results do not depend on when it runs)
    if (mode == 1) sc->pending = any && (nrun == 0 || 4 * nwait >= nwait + nrun || maxw >= 3);
    if (mode == 1) sc->certified = 0;
    if (mode == 3) sc->certified = cert;
    if (mode == 2 || mode == 3) sc->pending = 0;
    if (mode >= 1 && mode <= 3) sc->active = run;
    if (mode == 1) sc->restart = 0;
    if (mode == 3) sc->restart = rst;
  }
}

hipError_t launch_kb_reduce(hipStream_t st, int nparts, int k, const double* part, BScalars* sc, int mode,
                            const int* gate) {
  if (k > 64) return hipErrorInvalidValue;
  const int* g2 = mode == 3 ? &sc->pending : gate;
  hipLaunchKernelGGL(kb_colsum, dim3(k), dim3(256), 0, st, nparts, k, part, sc, g2);
  hipLaunchKernelGGL(kb_reduce, dim3(1), dim3(64), 0, st, k, sc, mode, gate);
  return hipGetLastError();
}

// H alpha of the columns certified by the last certification reduction (saved for
// the recovery); skipped on the device unless one was certified
__global__ __launch_bounds__(256) void kb_save_hg(int64_t n, int k, const BScalars* sc, const double* __restrict__ hg,
                                                  double* __restrict__ hgf) {
  if (sc->certified == 0) return;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += stride)
    if (sc->certnow[e % k]) hgf[e] = hg[e];
}

hipError_t launch_kb_save_hg(hipStream_t st, int n6, int k, BScalars* sc, const double* hg, double* hgf) {
  if (k > 64) return hipErrorInvalidValue;
  const int64_t n = (int64_t)n6 * k;
  hipLaunchKernelGGL(kb_save_hg, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0, st, n, k, sc,
                     hg, hgf);
  return hipGetLastError();
}

// mode 0: lam += rho w ; g += rho Fw      mode 1: w = Pz + gamma w      (running columns)
// mode 2: g = g_true, w = P M P g_true where repl[j]                     (restart)
// mode 3: a = c for joining columns (status 5)                            (start)
__global__ __launch_bounds__(256) void kb_update(int64_t nlam, int k, int mode, double* __restrict__ a,
                                                 double* __restrict__ b, const double* __restrict__ c,
                                                 const double* __restrict__ e, const BScalars* __restrict__ sc,
                                                 const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int64_t n = nlam * k;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int j = (int)(t % k);
    if (mode == 2) {
      if (sc->repl[j]) {
        a[t] = c[t];
        b[t] = e[t];
      }
    } else if (mode == 3) {
      if (sc->done[j] == 5) a[t] = c[t];
    } else if (!sc->done[j]) {
      if (mode == 0) {
        const double rho = sc->rho[j];
        a[t] += rho * c[t];
        b[t] += rho * e[t];
      } else {
        a[t] = c[t] + sc->gamma[j] * a[t];
      }
    }
  }
}

hipError_t launch_kb_update(hipStream_t st, int64_t nlam, int k, int mode, double* a, double* b, const double* c,
                            const double* e, const BScalars* sc, const int* gate) {
  const int64_t n = nlam * k;
  hipLaunchKernelGGL(kb_update, dim3((unsigned)std::min<int64_t>((n + 255) / 256, kStrideGrid)), dim3(256), 0, st, nlam,
                     k, mode, a, b, c, e, sc, gate);
  return hipGetLastError();
}

// packed local block (column j at j*npacked) <-> interior-first padded blocks
//   dir 0: Vr[(orrow[d] + i)*kp + j] = V[j][perm[i]] (i < nr), Vb[(obrow[d] + p)*kp + j] = V[j][perm[nr+p]]
//   dir 1: the reverse
__global__ __launch_bounds__(256) void kb_permute(int dir, int k, int kp, int64_t npacked, const int* __restrict__ nn,
                                                  const int* __restrict__ nr, const int* __restrict__ perm,
                                                  const int64_t* __restrict__ noff, double* __restrict__ V,
                                                  double* __restrict__ Vr, const int64_t* __restrict__ orrow,
                                                  double* __restrict__ Vb, const int64_t* __restrict__ obrow) {
  const int d = blockIdx.y;
  const int n = nn[d], r = nr[d];
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int i = (int)(t / k), j = (int)(t % k);
  if (i >= n) return;
  if (i < r && !Vr) return;  // interior rows left out
  const int l = perm[noff[d] + i];
  double* v = V + (int64_t)j * npacked + noff[d] + l;
  double* s = i < r ? Vr + (orrow[d] + i) * kp + j : Vb + (obrow[d] + (i - r)) * kp + j;
  if (dir == 0) *s = *v;
  else *v = *s;
}

hipError_t launch_kb_permute(hipStream_t st, int nsub, int maxn, int dir, int k, int kp, int64_t npacked,
                             const int* n, const int* nr, const int* perm, const int64_t* noff, double* V, double* Vr,
                             const int64_t* orrow, double* Vb, const int64_t* obrow) {
  const int64_t tot = (int64_t)maxn * k;
  hipLaunchKernelGGL(kb_permute, dim3((unsigned)((tot + 255) / 256), nsub), dim3(256), 0, st, dir, k, kp, npacked, n,
                     nr, perm, noff, V, Vr, orrow, Vb, obrow);
  return hipGetLastError();
}

// V[j][perm[nr+p]] -= (Bf^T lam)[p][j] at the interface DOFs of every subdomain
__global__ __launch_bounds__(256) void kb_sub_bt(int k, int64_t npacked, const int* __restrict__ nr,
                                                 const int* __restrict__ ns, const int* __restrict__ perm,
                                                 const int64_t* __restrict__ noff, const int64_t* __restrict__ boff,
                                                 const int* __restrict__ gptr, const int* __restrict__ gidx,
                                                 const double* __restrict__ lam, double* __restrict__ V) {
  const int d = blockIdx.y;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int p = (int)(t / k), j = (int)(t % k);
  if (p >= ns[d]) return;
  double s = 0.0;
  const int64_t q = boff[d] + p;
  for (int e = gptr[q]; e < gptr[q + 1]; e++) {
    const int code = gidx[e];
    const int r = code >= 0 ? code : ~code;
    s += code >= 0 ? lam[(int64_t)r * k + j] : -lam[(int64_t)r * k + j];
  }
  V[(int64_t)j * npacked + noff[d] + perm[noff[d] + nr[d] + p]] -= s;
}

hipError_t launch_kb_sub_bt(hipStream_t st, int nsub, int maxns, int k, int64_t npacked, const int* nr, const int* ns,
                            const int* perm, const int64_t* noff, const int64_t* boff, const int* gptr,
                            const int* gidx, const double* lam, double* V) {
  const int64_t tot = (int64_t)maxns * k;
  hipLaunchKernelGGL(kb_sub_bt, dim3((unsigned)((tot + 255) / 256), nsub), dim3(256), 0, st, k, npacked, nr, ns, perm,
                     noff, boff, gptr, gidx, lam, V);
  return hipGetLastError();
}

// U[j*nglob + gdof] = X[j][l] + R_i[l,:] (la[arow_d + m][j] + h[6d+m][j]) for owned DOFs
__global__ __launch_bounds__(256) void kb_recover(int k, int kp, int64_t npacked, int64_t nglob,
                                                  const int* __restrict__ nn, const int64_t* __restrict__ noff,
                                                  const double* __restrict__ R, const int64_t* __restrict__ roff,
                                                  const int64_t* __restrict__ arow, const double* __restrict__ la,
                                                  const double* __restrict__ h, const double* __restrict__ X,
                                                  const int* __restrict__ gdof, const unsigned char* __restrict__ owner,
                                                  double* __restrict__ U) {
  const int d = blockIdx.y;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int l = (int)(t / k), j = (int)(t % k);
  const int n = nn[d];
  if (l >= n) return;
  if (!owner[noff[d] + l]) return;
  const double* Rd = R + roff[d] + l;
  double s = X[(int64_t)j * npacked + noff[d] + l];
  for (int m = 0; m < 6; m++) s += Rd[(int64_t)m * n] * (la[(arow[d] + m) * kp + j] + h[(int64_t)(6 * d + m) * k + j]);
  U[(int64_t)j * nglob + gdof[noff[d] + l]] = s;
}

hipError_t launch_kb_recover(hipStream_t st, int nsub, int maxn, int k, int kp, int64_t npacked, int64_t nglob,
                             const int* n, const int64_t* noff, const double* R, const int64_t* roff,
                             const int64_t* arow, const double* la, const double* h, const double* X, const int* gdof,
                             const unsigned char* owner, double* U) {
  const int64_t tot = (int64_t)maxn * k;
  hipLaunchKernelGGL(kb_recover, dim3((unsigned)((tot + 255) / 256), nsub), dim3(256), 0, st, k, kp, npacked, nglob, n,
                     noff, R, roff, arow, la, h, X, gdof, owner, U);
  return hipGetLastError();
}

}  // namespace hfls
