// Batched tile GEMM on the matrix cores (gfx950 MFMA) for blocks of right-hand
// sides: Y_d = alpha * op(A_d) X_d (+ Z_d), one matrix per subdomain.  The FP32
// kernel accumulates in FP32 and stores FP64 (Y, Z double).
//
// A_d is stored as 64x64 tiles, column major inside a tile, zero padded:
//   mode 0 symmetric: lower tiles only, tile (i,j), i>=j, at j*T - j*(j-1)/2 + (i-j);
//          op(A)(i,j) = tile(i,j) for i>=j, tile(j,i)^T otherwise
//   mode 1 general, tile (i,j) at j*Ra + i; op 0: A, op 1: A^T
//   mode 2 compact column major with leading dimension lda[d] (no tiles); op 0/1
//   mode 3 lower triangular, stored like mode 0: op 0: A, op 1: A^T (FP64 only)
// X_d, Y_d, Z_d are row major (row = DOF, column = right-hand side) with leading
// dimension kp (a multiple of 64) and rows padded to whole tiles.  With a gate
// pointer the launch does nothing while *gate == 0 (device-side condition).
//
// One workgroup computes one 64-row x 64-column output block; the inner
// dimension is streamed in chunks of 32 through LDS (A chunk stored [k][row] with
// a padded row so both the transposed tile loads and the operand reads are free of
// bank conflicts), with 16-byte global loads of the next chunk issued before the
// matrix-core work on the current one.  FP32 takes tiles only (modes 0, 1).  FP32: v_mfma_f32_32x32x2f32, wave w owns the 32x32 block
// (w/2, w%2).  FP64: v_mfma_f64_16x16x4f64, wave w owns rows 16w..16w+15 and all
// four 16-column blocks.  Operand/result lane layouts (checked on gfx950):
//   f32 32x32x2: A[l%32][l/32], B[l/32][l%32], D row 8(r/4)+4(l/32)+r%4, col l%32
//   f64 16x16x4: A[l%16][l/16], B[l/16][l%16], D row 4r+l/16, col l%16
#include <algorithm>

#include "kernels.hpp"

namespace hfls {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

static __device__ __forceinline__ int64_t tidx_sym(int T, int i, int j) {
  return (int64_t)j * T - (int64_t)j * (j - 1) / 2 + (i - j);
}

// pointer to stored tile and whether it must be transposed, for op(A)(i,j)
static __device__ __forceinline__ int64_t tile_of(int mode, int op, int Ra, int i, int j, bool* tr) {
  if (mode == 3) {  // lower triangular: op 0 uses tile (i, j), j <= i; op 1 tile (j, i)^T, j >= i
    *tr = op == 1;
    return op == 0 ? tidx_sym(Ra, i, j) : tidx_sym(Ra, j, i);
  }
  if (mode == 0) {
    if (i >= j) {
      *tr = false;
      return tidx_sym(Ra, i, j);
    }
    *tr = true;
    return tidx_sym(Ra, j, i);
  }
  if (op == 0) {
    *tr = false;
    return (int64_t)j * Ra + i;
  }
  *tr = true;
  return (int64_t)i * Ra + j;  // stored tile (j, i)
}

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

// Per-thread staging of one K chunk (A: 32 x 64 of op(A)^T, X: 32 x 64): the next
// chunk is fetched into registers while the matrix cores work on the current one
// from LDS.  kind: 0 tile, 1 transposed tile, 2 compact op 0, 3 compact op 1.
struct StageF {
  float a[8], x[8];
};
struct StageD {
  double a[8], x[8];
};

static __device__ __forceinline__ void fetch_f32(const float* __restrict__ tile, int kind, int k0,
                                                 const float* __restrict__ Xrow, int kp, StageF& s) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int f = t + 256 * q;
    f4 v;
    if (kind == 0) v = *(const f4*)(tile + (k0 + (f >> 4)) * 64 + (f & 15) * 4);
    else v = *(const f4*)(tile + (f >> 3) * 64 + k0 + (f & 7) * 4);
    s.a[4 * q + 0] = v.x;
    s.a[4 * q + 1] = v.y;
    s.a[4 * q + 2] = v.z;
    s.a[4 * q + 3] = v.w;
    const f4 xv = *(const f4*)(Xrow + (int64_t)(k0 + (f >> 4)) * kp + (f & 15) * 4);
    s.x[4 * q + 0] = xv.x;
    s.x[4 * q + 1] = xv.y;
    s.x[4 * q + 2] = xv.z;
    s.x[4 * q + 3] = xv.w;
  }
}

static __device__ __forceinline__ void stash_f32(int kind, const StageF& s, float* As, float* Xs) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int f = t + 256 * q;
    if (kind == 0) {
      const int kk = f >> 4, r4 = (f & 15) * 4;
#pragma unroll
      for (int u = 0; u < 4; u++) As[kk * 65 + r4 + u] = s.a[4 * q + u];
    } else {
      const int row = f >> 3, k4 = (f & 7) * 4;
#pragma unroll
      for (int u = 0; u < 4; u++) As[(k4 + u) * 65 + row] = s.a[4 * q + u];
    }
    f4 xv;
    xv.x = s.x[4 * q + 0];
    xv.y = s.x[4 * q + 1];
    xv.z = s.x[4 * q + 2];
    xv.w = s.x[4 * q + 3];
    *(f4*)(Xs + (f >> 4) * 64 + (f & 15) * 4) = xv;
  }
}

template <bool CM>
static __device__ __forceinline__ void fetch_f64(const double* __restrict__ tile, const double* __restrict__ Ac,
                                                 int64_t lda, int m, int n, int i, int j, int kind, int k0,
                                                 const double* __restrict__ Xrow, int kp, StageD& s) {
  const int t = threadIdx.x;
  if (!CM) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int f = t + 256 * q;
      d2 v;
      if (kind == 0) v = *(const d2*)(tile + (k0 + (f >> 5)) * 64 + (f & 31) * 2);
      else v = *(const d2*)(tile + (f >> 4) * 64 + k0 + (f & 15) * 2);
      s.a[2 * q + 0] = v.x;
      s.a[2 * q + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = t + 256 * q;
      if (kind == 2) {
        const int r = 64 * i + (e & 63), c = 64 * j + k0 + (e >> 6);
        s.a[q] = (r < m && c < n) ? Ac[r + (int64_t)c * lda] : 0.0;
      } else {
        const int r = 64 * i + (e >> 5), c = 64 * j + k0 + (e & 31);  // op(A)(r, c) = A(c, r)
        s.a[q] = (c < m && r < n) ? Ac[c + (int64_t)r * lda] : 0.0;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int f = t + 256 * q;
    const d2 xv = *(const d2*)(Xrow + (int64_t)(k0 + (f >> 5)) * kp + (f & 31) * 2);
    s.x[2 * q + 0] = xv.x;
    s.x[2 * q + 1] = xv.y;
  }
}

template <bool CM>
static __device__ __forceinline__ void stash_f64(int kind, const StageD& s, double* As, double* Xs) {
  const int t = threadIdx.x;
  if (!CM) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int f = t + 256 * q;
      if (kind == 0) {
        const int kk = f >> 5, r2 = (f & 31) * 2;
        As[kk * 65 + r2] = s.a[2 * q];
        As[kk * 65 + r2 + 1] = s.a[2 * q + 1];
      } else {
        const int row = f >> 4, k2 = (f & 15) * 2;
        As[k2 * 65 + row] = s.a[2 * q];
        As[(k2 + 1) * 65 + row] = s.a[2 * q + 1];
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = t + 256 * q;
      if (kind == 2) As[(e >> 6) * 65 + (e & 63)] = s.a[q];
      else As[(e & 31) * 65 + (e >> 5)] = s.a[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int f = t + 256 * q;
    d2 xv;
    xv.x = s.x[2 * q];
    xv.y = s.x[2 * q + 1];
    *(d2*)(Xs + (f >> 5) * 64 + (f & 31) * 2) = xv;
  }
}

// source of K chunk c (tile j = c/2, rows k0 = 32*(c%2) of it)
static __device__ __forceinline__ int chunk_kind(int mode, int op, int Ra, int i, int c, int64_t* ti) {
  bool tr = false;
  *ti = mode == 2 ? 0 : tile_of(mode, op, Ra, i, c >> 1, &tr);
  if (mode == 2) return op == 0 ? 2 : 3;
  return tr ? 1 : 0;
}

// RT row tiles (64 rows each) per workgroup share the X chunk in LDS: RT = 2
// halves the X staging and LDS reads per matrix-core operation (chosen per mesh
// by the autotuner of the block solve).
//
// Workgroup order: the hardware hands consecutive workgroups to the 8 XCDs in
// turn, each XCD with its own L2.  The linear workgroup id is mapped so that all
// row blocks of one subdomain run on one XCD (slot s = id / 8 of XCD id % 8 is row
// block s % nslot of the subdomain at entry 8 (s / nslot) + id % 8 of the XCD
// order, or of subdomain 8 (s / nslot) + id % 8 without one), and row block i takes the
// tile columns in the order j = (t - r_i) mod Ko, t = 0, 1, ..., with r_i = i & ~1
// (the even row block of its pair): at step t, row blocks i and j = t - r_i need
// the same symmetric tile (max(i,j), min(i,j)) within a step of each other, one
// directly, one transposed.  The X block of the subdomain is served by the XCD's
// L2; the second tile read mostly is not (FETCH_SIZE is about twice the packed
// tiles: the workgroups drift apart by more than the L2 holds).  Both tile shapes
// (RT = 1, 2) use the same r_i, so they add every row's products in the same order
// and give bit-identical results.
template <int RT>
__global__ __launch_bounds__(256) void k_tgemm_f32(int mode, int op, int nsub, int nslot,
                                                   const int* __restrict__ order, int norder,
                                                   const int* __restrict__ mrows,
                                                   const int* __restrict__ mcols, const float* __restrict__ A,
                                                   const int64_t* __restrict__ aoff, const float* __restrict__ X,
                                                   const int64_t* __restrict__ xoff, double* __restrict__ Y,
                                                   const int64_t* __restrict__ yoff, const double* __restrict__ Z,
                                                   int kp, float alpha, const int* __restrict__ gate,
                                                   float* __restrict__ Yf) {
  const int sl = blockIdx.x >> 3;
  const int gi = (sl / nslot) * 8 + (blockIdx.x & 7);
  const int d = order ? (gi < norder ? order[gi] : -1) : (gi < nsub ? gi : -1);
  if (d < 0) return;
  if (gate && *gate == 0) return;
  const int Ra = (mrows[d] + 63) / 64, Ca = (mcols[d] + 63) / 64;
  const int Ro = op == 1 && mode != 0 ? Ca : Ra;
  const int Ko = op == 1 && mode != 0 ? Ra : Ca;
  const int i0 = (sl % nslot) * RT;
  if (i0 >= Ro) return;
  const int rot = (i0 & ~1) % Ko;  // K chunk c covers tile column (c/2 - rot) mod Ko
  const int nrt = min(RT, Ro - i0);
  const int c0 = blockIdx.z * 64;
  __shared__ float As[RT][32 * 65];
  __shared__ float Xs[32 * 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rb = (w >> 1) * 32, cb = (w & 1) * 32;
  floatx16 acc[RT];
#pragma unroll
  for (int q = 0; q < RT; q++)
#pragma unroll
    for (int r = 0; r < 16; r++) acc[q][r] = 0.0f;
  const float* Ad = A + aoff[d];
  const float* Xd = X + xoff[d] + c0;
  const int nch = 2 * Ko;
  StageF s[RT];
  int kind[RT];
  auto fetch = [&](int c) {
    int tj = (c >> 1) - rot;
    if (tj < 0) tj += Ko;
    const int cc = 2 * tj + (c & 1);
#pragma unroll
    for (int q = 0; q < RT; q++) {
      if (q < nrt) {
        int64_t ti;
        kind[q] = chunk_kind(mode, op, Ra, i0 + q, cc, &ti);
        fetch_f32(Ad + ti * 4096, kind[q], (c & 1) * 32, Xd + (int64_t)(tj * 64) * kp, kp, s[q]);
      }
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int q = 0; q < RT; q++)
      if (q < nrt) stash_f32(kind[q], s[q], As[q], Xs);
  };
  if (nch > 0) {
    fetch(0);
    stash();
  }
  __syncthreads();
  for (int c = 0; c < nch; c++) {
    const bool more = c + 1 < nch;
    if (more) fetch(c + 1);
    __builtin_amdgcn_iglp_opt(0);  // interleave the LDS operand reads with the MFMAs
#pragma unroll
    for (int kk = 0; kk < 32; kk += 2) {
      const float b = Xs[(kk + (lane >> 5)) * 64 + cb + (lane & 31)];
#pragma unroll
      for (int q = 0; q < RT; q++) {
        const float a = As[q][(kk + (lane >> 5)) * 65 + rb + (lane & 31)];
        acc[q] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[q], 0, 0, 0);
      }
    }
    __syncthreads();
    if (more) stash();
    __syncthreads();
  }
  if (Yf) {  // FP32 result: half the store traffic, the same values
    float* Yfd = Yf + yoff[d] + c0;
#pragma unroll
    for (int q = 0; q < RT; q++) {
      if (q >= nrt) break;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int row = (i0 + q) * 64 + rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        Yfd[(int64_t)row * kp + cb + (lane & 31)] = alpha * acc[q][r];
      }
    }
    return;
  }
  double* Yd = Y + yoff[d] + c0;
  const double* Zd = Z ? Z + yoff[d] + c0 : nullptr;
#pragma unroll
  for (int q = 0; q < RT; q++) {
    if (q >= nrt) break;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int row = (i0 + q) * 64 + rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int col = cb + (lane & 31);
      double v = (double)(alpha * acc[q][r]);
      if (Zd) v += Zd[(int64_t)row * kp + col];
      Yd[(int64_t)row * kp + col] = v;
    }
  }
}

// S > 1 (split K): slot q of a matrix is row block i = q / S and K part s = q % S;
// the workgroup takes K chunks [s*nch/S, (s+1)*nch/S) and writes its partial
// 64 x kp block to Ypart at ((d*Rmax + i)*S + s)*64*kp (Rmax = nslot / S, the
// largest row-block count); k_splitk_sum adds the S partials in a fixed order.
// Workgroups are grouped per XCD as in k_tgemm_f32 (nslot = row blocks, or S);
// with symmetric tiles and no split, row block i takes the tile columns in the
// order (t - i) mod Ko, so the two reads of every off-diagonal tile meet in L2.
template <bool CM>  // compact column-major A (mode 2) or tiles
__global__ __launch_bounds__(256) void k_tgemm_f64(int mode, int op, int nsub, int nslot,
                                                   const int* __restrict__ order, int norder,
                                                   const int* __restrict__ mrows,
                                                   const int* __restrict__ mcols, const double* __restrict__ A,
                                                   const int64_t* __restrict__ aoff, const int* __restrict__ lda,
                                                   const double* __restrict__ X,
                                                   const int64_t* __restrict__ xoff, double* __restrict__ Y,
                                                   const int64_t* __restrict__ yoff, const double* __restrict__ Z,
                                                   int kp, double alpha, const int* __restrict__ gate, int S,
                                                   double* __restrict__ Ypart, const float* __restrict__ Zf) {
  const int sl = blockIdx.x >> 3;
  const int gi = (sl / nslot) * 8 + (blockIdx.x & 7);
  const int d = order ? (gi < norder ? order[gi] : -1) : (gi < nsub ? gi : -1);
  if (d < 0) return;
  if (gate && *gate == 0) return;
  const int m = mrows[d], n = mcols[d];
  const int Ra = (m + 63) / 64, Ca = (n + 63) / 64;
  const int Ro = (mode != 0 && op == 1) ? Ca : Ra;
  const int Ko = (mode != 0 && op == 1) ? Ra : Ca;
  const int i = S > 1 ? (sl % nslot) / S : sl % nslot;
  const int sp = S > 1 ? (sl % nslot) % S : 0;
  if (i >= Ro) return;
  const int rot = (mode == 0 && S == 1) ? i : 0;  // K chunk c covers tile column (c/2 - rot) mod Ko
  const int c0 = blockIdx.z * 64;
  __shared__ double As[32 * 65];
  __shared__ double Xs[32 * 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rb = w * 16;
  doublex4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[c][r] = 0.0;
  const double* Ad = A + aoff[d];
  const int64_t la = mode == 2 ? (int64_t)lda[d] : 0;
  const double* Xd = X + xoff[d] + c0;
  const int nall = 2 * Ko;
  int cbeg = (int)((int64_t)sp * nall / S), cend = (int)((int64_t)(sp + 1) * nall / S);
  if (mode == 3) {  // only the tiles of the triangle (S == 1)
    cbeg = op == 0 ? 0 : 2 * i;
    cend = op == 0 ? 2 * (i + 1) : nall;
  }
  StageD s;
  auto fetch = [&](int c) {
    int tj = (c >> 1) - rot;
    if (tj < 0) tj += Ko;
    int64_t ti;
    const int kind = chunk_kind(mode, op, Ra, i, 2 * tj + (c & 1), &ti);
    fetch_f64<CM>(Ad + ti * 4096, Ad, la, m, n, i, tj, kind, (c & 1) * 32, Xd + (int64_t)(tj * 64) * kp, kp, s);
    return kind;
  };
  if (cend > cbeg) stash_f64<CM>(fetch(cbeg), s, As, Xs);
  __syncthreads();
  for (int c = cbeg; c < cend; c++) {
    const bool more = c + 1 < cend;
    int kind_n = 0;
    if (more) kind_n = fetch(c + 1);
#pragma unroll
    for (int kk = 0; kk < 32; kk += 4) {
      const double a = As[(kk + (lane >> 4)) * 65 + rb + (lane & 15)];
#pragma unroll
      for (int cc = 0; cc < 4; cc++) {
        const double b = Xs[(kk + (lane >> 4)) * 64 + 16 * cc + (lane & 15)];
        acc[cc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[cc], 0, 0, 0);
      }
    }
    __syncthreads();
    if (more) stash_f64<CM>(kind_n, s, As, Xs);
    __syncthreads();
  }
  if (S > 1) {
    double* Pd = Ypart + (((int64_t)d * (nslot / S) + i) * S + sp) * 64 * kp + c0;
#pragma unroll
    for (int cc = 0; cc < 4; cc++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        Pd[(int64_t)(rb + 4 * r + (lane >> 4)) * kp + 16 * cc + (lane & 15)] = alpha * acc[cc][r];
    return;
  }
  double* Yd = Y + yoff[d] + c0;
  const double* Zd = Z ? Z + yoff[d] + c0 : nullptr;
  const float* Zfd = Zf ? Zf + yoff[d] + c0 : nullptr;
#pragma unroll
  for (int cc = 0; cc < 4; cc++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = i * 64 + rb + 4 * r + (lane >> 4);
      const int col = 16 * cc + (lane & 15);
      double v = alpha * acc[cc][r];
      if (Zfd) v += (double)Zfd[(int64_t)row * kp + col];
      else if (Zd) v += Zd[(int64_t)row * kp + col];
      Yd[(int64_t)row * kp + col] = v;
    }
}

// Y[yoff[d] + (64 i + r)*kp + j] = sum_{s < S} Ypart[(((d*Rmax + i)*S + s)*64 + r)*kp + j]
// (+ Z), fixed order, for the row blocks i < ceil(orows[d] / 64)
__global__ __launch_bounds__(256) void k_splitk_sum(int S, int Rmax, int kp, const int* __restrict__ orows,
                                                    const double* __restrict__ Ypart, double* __restrict__ Y,
                                                    const int64_t* __restrict__ yoff, const double* __restrict__ Z,
                                                    const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int d = blockIdx.y / Rmax, i = blockIdx.y % Rmax;
  if (i >= (orows[d] + 63) / 64) return;
  const int e = blockIdx.x * 256 + threadIdx.x;  // 64 * kp entries
  if (e >= 64 * kp) return;
  const double* p = Ypart + ((int64_t)d * Rmax + i) * S * 64 * kp + e;
  double v = 0.0;
  for (int s = 0; s < S; s++) v += p[(int64_t)s * 64 * kp];
  const int64_t o = yoff[d] + (int64_t)i * 64 * kp + e;
  if (Z) v += Z[o];
  Y[o] = v;
}

hipError_t launch_tgemm_f32(hipStream_t st, int nsub, int maxRo, int mode, int op, const int* mrows, const int* mcols,
                            const float* A, const int64_t* aoff, const int* lda, const float* X, const int64_t* xoff, double* Y,
                            const int64_t* yoff, const double* Z, int kp, float alpha, const int* gate, int rt,
                            float* Yf, const int* order, int norder) {
  (void)lda;
  if (kp % 64 || mode == 2 || mode == 3) return hipErrorInvalidValue;  // FP32: modes 0, 1
  if (order && (norder < 8 || norder % 8)) return hipErrorInvalidValue;
  // subdomain groups of 8, one subdomain of each per XCD
  const int ngroups = order ? norder / 8 : (nsub + 7) / 8;
  if (rt == 2) {
    const int nslot = (maxRo + 1) / 2;
    hipLaunchKernelGGL(k_tgemm_f32<2>, dim3(8 * nslot * ngroups, 1, kp / 64), dim3(256), 0, st, mode, op, nsub, nslot,
                       order, norder, mrows, mcols, A, aoff, X, xoff, Y, yoff, Z, kp, alpha, gate, Yf);
  } else {
    hipLaunchKernelGGL(k_tgemm_f32<1>, dim3(8 * maxRo * ngroups, 1, kp / 64), dim3(256), 0, st, mode, op, nsub, maxRo,
                       order, norder, mrows, mcols, A, aoff, X, xoff, Y, yoff, Z, kp, alpha, gate, Yf);
  }
  return hipGetLastError();
}

hipError_t launch_tgemm_f64(hipStream_t st, int nsub, int maxRo, int mode, int op, const int* mrows, const int* mcols,
                            const double* A, const int64_t* aoff, const int* lda, const double* X, const int64_t* xoff, double* Y,
                            const int64_t* yoff, const double* Z, int kp, double alpha, const int* gate, int S,
                            double* Ypart, const float* Zf, const int* order, int norder) {
  if (kp % 64) return hipErrorInvalidValue;
  if (mode == 2 && !lda) return hipErrorInvalidValue;
  if (S > 1 && (!Ypart || mode == 3 || Zf)) return hipErrorInvalidValue;
  if (order && (norder < 8 || norder % 8)) return hipErrorInvalidValue;
  const int nslot = S > 1 ? S * maxRo : maxRo;
  const dim3 grid(8 * nslot * (order ? norder / 8 : (nsub + 7) / 8), 1, kp / 64);
  if (mode == 2)
    hipLaunchKernelGGL(k_tgemm_f64<true>, grid, dim3(256), 0, st, mode, op, nsub, nslot, order, norder, mrows, mcols, A,
                       aoff, lda, X, xoff, Y, yoff, Z, kp, alpha, gate, S, Ypart, Zf);
  else
    hipLaunchKernelGGL(k_tgemm_f64<false>, grid, dim3(256), 0, st, mode, op, nsub, nslot, order, norder, mrows, mcols,
                       A, aoff, lda, X, xoff, Y, yoff, Z, kp, alpha, gate, S, Ypart, Zf);
  if (S > 1)  // output rows: mrows, or mcols for the transpose of a general or compact A
    hipLaunchKernelGGL(k_splitk_sum, dim3((64 * kp + 255) / 256, nsub * maxRo), dim3(256), 0, st, S, maxRo, kp,
                       (mode != 0 && op == 1) ? mcols : mrows, Ypart, Y, yoff, Z, gate);
  return hipGetLastError();
}

// Y_d = A_d C_d (+ Z_d or + Zf_d) for a general tiled A_d with at most 64 columns (one
// tile column: tile i at A + aoff[d] + 4096 i) and C_d of as many rows (row major, ld
// kp, C + coff[d]): the HTFETI correction y_b += [C|R_b|Q|Yh] coef of the block solve.
// One workgroup per 64-row output block; C_d (the same for every row block of the
// subdomain) is staged once in LDS, rows past mcols[d] as zeros; every wave takes 16
// output rows and loads its A operands straight from the column-major tile (one
// 128-byte row segment per 16 lanes and K step), all of them before the first matrix
// core operation, so the workgroup has one barrier.  Same workgroup order and
// results as k_tgemm_f64 mode 1 op 0 up to the summation order inside a 64-wide K.
__global__ __launch_bounds__(256) void k_corr_f64(int nsub, int nslot, const int* __restrict__ order, int norder,
                                                  const int* __restrict__ mrows, const int* __restrict__ mcols,
                                                  const double* __restrict__ A, const int64_t* __restrict__ aoff,
                                                  const double* __restrict__ C, const int64_t* __restrict__ coff,
                                                  double* __restrict__ Y, const int64_t* __restrict__ yoff,
                                                  const double* __restrict__ Z, const float* __restrict__ Zf, int kp,
                                                  const int* __restrict__ gate) {
  const int sl = blockIdx.x >> 3;
  const int gi = (sl / nslot) * 8 + (blockIdx.x & 7);
  const int d = order ? (gi < norder ? order[gi] : -1) : (gi < nsub ? gi : -1);
  if (d < 0) return;
  if (gate && *gate == 0) return;
  const int m = mrows[d], nc = mcols[d];
  const int i = sl % nslot;
  if (i >= (m + 63) / 64) return;
  const int c0 = blockIdx.z * 64;
  __shared__ double Cs[64 * 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, rb = 16 * w;
  const double* Ai = A + aoff[d] + (int64_t)i * 4096 + rb + (lane & 15);
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; q++) a[q] = Ai[(4 * q + (lane >> 4)) * 64];  // A[rb + l%16][4q + l/16]
  const double* Cd = C + coff[d] + c0;
  for (int t = threadIdx.x; t < 64 * 64; t += 256) {
    const int r = t >> 6, c = t & 63;
    Cs[t] = r < nc ? Cd[(int64_t)r * kp + c] : 0.0;
  }
  __syncthreads();
  doublex4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) acc[c][r] = 0.0;
#pragma unroll
  for (int q = 0; q < 16; q++)
#pragma unroll
    for (int cc = 0; cc < 4; cc++)
      acc[cc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], Cs[(4 * q + (lane >> 4)) * 64 + 16 * cc + (lane & 15)],
                                                     acc[cc], 0, 0, 0);
  double* Yd = Y + yoff[d] + c0;
  const double* Zd = Z ? Z + yoff[d] + c0 : nullptr;
  const float* Zfd = Zf ? Zf + yoff[d] + c0 : nullptr;
#pragma unroll
  for (int cc = 0; cc < 4; cc++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = i * 64 + rb + 4 * r + (lane >> 4);
      const int col = 16 * cc + (lane & 15);
      double v = acc[cc][r];
      if (Zfd) v += (double)Zfd[(int64_t)row * kp + col];
      else if (Zd) v += Zd[(int64_t)row * kp + col];
      Yd[(int64_t)row * kp + col] = v;
    }
}

hipError_t launch_corr_f64(hipStream_t st, int nsub, int maxRo, const int* mrows, const int* mcols, const double* A,
                           const int64_t* aoff, const double* C, const int64_t* coff, double* Y, const int64_t* yoff,
                           const double* Z, const float* Zf, int kp, const int* gate, const int* order, int norder) {
  if (kp % 64) return hipErrorInvalidValue;
  if (order && (norder < 8 || norder % 8)) return hipErrorInvalidValue;
  const dim3 grid(8 * maxRo * (order ? norder / 8 : (nsub + 7) / 8), 1, kp / 64);
  hipLaunchKernelGGL(k_corr_f64, grid, dim3(256), 0, st, nsub, maxRo, order, norder, mrows, mcols, A, aoff, C, coff, Y,
                     yoff, Z, Zf, kp, gate);
  return hipGetLastError();
}

// pack a compact column-major matrix (m x n, ld) into 64x64 tiles (mode 0: lower
// tiles of a symmetric matrix, mode 1: all tiles), converting to T
template <typename T>
__global__ __launch_bounds__(256) void k_pack_any(int mode, const int* __restrict__ mrows, const int* __restrict__ mcols,
                                                  const double* __restrict__ A, const int64_t* __restrict__ aoff,
                                                  const int* __restrict__ lda, T* __restrict__ P,
                                                  const int64_t* __restrict__ poff) {
  const int d = blockIdx.y;
  const int m = mrows[d], n = mcols[d];
  const int Ra = (m + 63) / 64, Ca = (n + 63) / 64;
  const int q = blockIdx.x;
  int ti, tj;
  int64_t idx;
  if (mode == 0) {
    if (q >= Ra * (Ra + 1) / 2) return;
    tj = 0;
    int rem = q;
    while (rem >= Ra - tj) {
      rem -= Ra - tj;
      tj++;
    }
    ti = tj + rem;
    idx = tidx_sym(Ra, ti, tj);
  } else {
    if (q >= Ra * Ca) return;
    ti = q % Ra;
    tj = q / Ra;
    idx = (int64_t)tj * Ra + ti;
  }
  const double* Ad = A + aoff[d];
  const int64_t l = lda[d];
  T* tile = P + poff[d] + idx * 4096;
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int c = e >> 6, r = e & 63;
    const int gr = 64 * ti + r, gc = 64 * tj + c;
    tile[e] = (gr < m && gc < n) ? (T)Ad[gr + (int64_t)gc * l] : (T)0;
  }
}

hipError_t launch_pack_f32(hipStream_t st, int nsub, int maxtiles, int mode, const int* mrows, const int* mcols,
                           const double* A, const int64_t* aoff, const int* lda, float* P, const int64_t* poff) {
  hipLaunchKernelGGL(k_pack_any<float>, dim3(maxtiles, nsub), dim3(256), 0, st, mode, mrows, mcols, A, aoff, lda, P,
                     poff);
  return hipGetLastError();
}

hipError_t launch_pack_f64(hipStream_t st, int nsub, int maxtiles, int mode, const int* mrows, const int* mcols,
                           const double* A, const int64_t* aoff, const int* lda, double* P, const int64_t* poff) {
  hipLaunchKernelGGL(k_pack_any<double>, dim3(maxtiles, nsub), dim3(256), 0, st, mode, mrows, mcols, A, aoff, lda, P,
                     poff);
  return hipGetLastError();
}

// matrix-core ceiling: 8 (f64) or 4 (f32) independent accumulator chains per wave,
// operands in registers drawn per round from a xorshift stream (uniform in
// [-0.5, 0.5) with random mantissas, so the datapaths toggle as on real data), so
// the launch measures the sustained matrix-core rate under the clock and power the
// device holds for it
static __device__ __forceinline__ unsigned xorshift(unsigned h) {
  h ^= h << 13;
  h ^= h >> 17;
  h ^= h << 5;
  return h;
}
static __device__ __forceinline__ float unit_half(unsigned bits) {
  return __uint_as_float((bits & 0x007fffffu) | 0x3f800000u) - 1.5f;
}

__global__ __launch_bounds__(256) void k_mfma_ceiling_f64(double* __restrict__ out, int iters, unsigned seed) {
  unsigned h = seed ^ (threadIdx.x * 2654435761u) ^ (blockIdx.x * 40503u);
  doublex4 c[8];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) c[i][r] = 0.0;
  for (int it = 0; it < iters; it++) {
    h = xorshift(h);
    const double a = unit_half(h), b = unit_half(h >> 9);
#pragma unroll
    for (int i = 0; i < 8; i++) c[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[i], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int r = 0; r < 4; r++) s += c[i][r];
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mfma_ceiling_f32(double* __restrict__ out, int iters, unsigned seed) {
  unsigned h = seed ^ (threadIdx.x * 2654435761u) ^ (blockIdx.x * 40503u);
  floatx16 c[4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int r = 0; r < 16; r++) c[i][r] = 0.0f;
  for (int it = 0; it < iters; it++) {
    h = xorshift(h);
    const float a = unit_half(h), b = unit_half(h >> 9);
#pragma unroll
    for (int i = 0; i < 4; i++) c[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c[i], 0, 0, 0);
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int r = 0; r < 16; r++) s += c[i][r];
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

std::vector<int> xcd_order(const std::vector<double>& work) {
  const int n = (int)work.size();
  std::vector<int> idx(n);
  for (int d = 0; d < n; d++) idx[d] = d;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return work[a] > work[b]; });
  std::vector<std::vector<int>> per(8);
  double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int d : idx) {
    int x = 0;
    for (int y = 1; y < 8; y++)
      if (load[y] < load[x]) x = y;
    per[x].push_back(d);
    load[x] += work[d];
  }
  size_t g = 0;
  for (auto& v : per) g = std::max(g, v.size());
  std::vector<int> out(8 * std::max<size_t>(g, 1), -1);
  for (int x = 0; x < 8; x++)
    for (size_t i = 0; i < per[x].size(); i++) out[8 * i + x] = per[x][i];
  return out;
}

hipError_t launch_mfma_ceiling(hipStream_t st, bool f64, int nwg, int iters, double* out, double* flops) {
  if (nwg < 1 || iters < 1) return hipErrorInvalidValue;
  const double waves = 4.0 * nwg;
  if (f64) {
    hipLaunchKernelGGL(k_mfma_ceiling_f64, dim3(nwg), dim3(256), 0, st, out, iters, 12345u);
    if (flops) *flops = waves * iters * 8 * (2.0 * 16 * 16 * 4);
  } else {
    hipLaunchKernelGGL(k_mfma_ceiling_f32, dim3(nwg), dim3(256), 0, st, out, iters, 12345u);
    if (flops) *flops = waves * iters * 4 * (2.0 * 32 * 32 * 2);
  }
  return hipGetLastError();
}

}  // namespace hfls
