// Batched tile GEMM on the matrix cores (gfx950 MFMA) for blocks of right-hand
// sides: Y_d = alpha * op(A_d) X_d (+ Z_d), one matrix per subdomain.
//
// A_d is stored as 64x64 tiles, column major inside a tile, zero padded:
//   mode 0 symmetric: lower tiles only, tile (i,j), i>=j, at j*T - j*(j-1)/2 + (i-j);
//          op(A)(i,j) = tile(i,j) for i>=j, tile(j,i)^T otherwise
//   mode 1 general, tile (i,j) at j*Ra + i; op 0: A, op 1: A^T
//   mode 2 compact column major with leading dimension lda[d] (no tiles); op 0/1
// X_d, Y_d, Z_d are row major (row = DOF, column = right-hand side) with leading
// dimension kp (a multiple of 64) and rows padded to whole tiles.  With a gate
// pointer the launch does nothing while *gate == 0 (device-side condition).
//
// One workgroup computes one 64-row x 64-column output block; the inner
// dimension is streamed in chunks of 32 through LDS (A chunk stored [k][row] with
// a padded row so both the transposed tile loads and the operand reads are free of
// bank conflicts).  FP32: v_mfma_f32_32x32x2f32, wave w owns the 32x32 block
// (w/2, w%2).  FP64: v_mfma_f64_16x16x4f64, wave w owns rows 16w..16w+15 and all
// four 16-column blocks.  Operand/result lane layouts (checked on gfx950):
//   f32 32x32x2: A[l%32][l/32], B[l/32][l%32], D row 8(r/4)+4(l/32)+r%4, col l%32
//   f64 16x16x4: A[l%16][l/16], B[l/16][l%16], D row 4r+l/16, col l%16
#include "kernels.hpp"

namespace hfls {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

static __device__ __forceinline__ int64_t tidx_sym(int T, int i, int j) {
  return (int64_t)j * T - (int64_t)j * (j - 1) / 2 + (i - j);
}

// pointer to stored tile and whether it must be transposed, for op(A)(i,j)
static __device__ __forceinline__ int64_t tile_of(int mode, int op, int Ra, int i, int j, bool* tr) {
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

template <typename T>
struct TG;

template <>
struct TG<float> {
  static constexpr int LDA = 65;
};
template <>
struct TG<double> {
  static constexpr int LDA = 65;
};

template <typename T>
static __device__ __forceinline__ void load_chunk(const T* __restrict__ tile, bool tr, int k0, T* As,
                                                  const T* __restrict__ Xrow, int kp, T* Xs) {
  const int t = threadIdx.x;
  if (!tr) {
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int e = t + 256 * m;  // 32 x 64
      const int kk = e >> 6, row = e & 63;
      As[kk * 65 + row] = tile[(k0 + kk) * 64 + row];
    }
  } else {
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int e = t + 256 * m;
      const int row = e >> 5, kk = e & 31;
      As[kk * 65 + row] = tile[row * 64 + k0 + kk];
    }
  }
#pragma unroll
  for (int m = 0; m < 8; m++) {
    const int e = t + 256 * m;
    const int kk = e >> 6, col = e & 63;
    Xs[kk * 64 + col] = Xrow[(int64_t)(k0 + kk) * kp + col];
  }
}

// chunk loader for a compact column-major A (m x n, lda): op(A) rows 64i.., inner 64j+k0..
template <typename T>
static __device__ __forceinline__ void load_chunk_cm(const T* __restrict__ A, int64_t lda, int m, int n, int op,
                                                     int i, int j, int k0, T* As, const T* __restrict__ Xrow, int kp,
                                                     T* Xs) {
  const int t = threadIdx.x;
  if (op == 0) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = t + 256 * q;
      const int kk = e >> 6, row = e & 63;
      const int r = 64 * i + row, c = 64 * j + k0 + kk;
      As[kk * 65 + row] = (r < m && c < n) ? A[r + (int64_t)c * lda] : (T)0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = t + 256 * q;
      const int row = e >> 5, kk = e & 31;
      const int r = 64 * i + row, c = 64 * j + k0 + kk;  // op(A)(r, c) = A(c, r)
      As[kk * 65 + row] = (c < m && r < n) ? A[c + (int64_t)r * lda] : (T)0;
    }
  }
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = t + 256 * q;
    const int kk = e >> 6, col = e & 63;
    Xs[kk * 64 + col] = Xrow[(int64_t)(k0 + kk) * kp + col];
  }
}

__global__ __launch_bounds__(256) void k_tgemm_f32(int mode, int op, const int* __restrict__ mrows,
                                                   const int* __restrict__ mcols, const float* __restrict__ A,
                                                   const int64_t* __restrict__ aoff, const int* __restrict__ lda,
                                                   const float* __restrict__ X,
                                                   const int64_t* __restrict__ xoff, float* __restrict__ Y,
                                                   const int64_t* __restrict__ yoff, const float* __restrict__ Z,
                                                   int kp, float alpha, const int* __restrict__ gate) {
  const int d = blockIdx.y;
  if (gate && *gate == 0) return;
  const int Ra = (mrows[d] + 63) / 64, Ca = (mcols[d] + 63) / 64;
  const int Ro = (mode != 0 && op == 1) ? Ca : Ra;
  const int Ko = (mode != 0 && op == 1) ? Ra : Ca;
  const int i = blockIdx.x;
  if (i >= Ro) return;
  const int c0 = blockIdx.z * 64;
  __shared__ float As[32 * 65];
  __shared__ float Xs[32 * 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rb = (w >> 1) * 32, cb = (w & 1) * 32;
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  const float* Ad = A + aoff[d];
  const float* Xd = X + xoff[d] + c0;
  for (int j = 0; j < Ko; j++) {
    bool tr = false;
    const int64_t ti = mode == 2 ? 0 : tile_of(mode, op, Ra, i, j, &tr);
    const float* tile = Ad + ti * 4096;
    for (int k0 = 0; k0 < 64; k0 += 32) {
      if (mode == 2)
        load_chunk_cm<float>(Ad, lda[d], mrows[d], mcols[d], op, i, j, k0, As, Xd + (int64_t)(j * 64) * kp, kp, Xs);
      else
        load_chunk<float>(tile, tr, k0, As, Xd + (int64_t)(j * 64) * kp, kp, Xs);
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < 32; kk += 2) {
        const float a = As[(kk + (lane >> 5)) * 65 + rb + (lane & 31)];
        const float b = Xs[(kk + (lane >> 5)) * 64 + cb + (lane & 31)];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
      }
      __syncthreads();
    }
  }
  float* Yd = Y + yoff[d] + c0;
  const float* Zd = Z ? Z + yoff[d] + c0 : nullptr;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int row = i * 64 + rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int col = cb + (lane & 31);
    float v = alpha * acc[r];
    if (Zd) v += Zd[(int64_t)row * kp + col];
    Yd[(int64_t)row * kp + col] = v;
  }
}

__global__ __launch_bounds__(256) void k_tgemm_f64(int mode, int op, const int* __restrict__ mrows,
                                                   const int* __restrict__ mcols, const double* __restrict__ A,
                                                   const int64_t* __restrict__ aoff, const int* __restrict__ lda,
                                                   const double* __restrict__ X,
                                                   const int64_t* __restrict__ xoff, double* __restrict__ Y,
                                                   const int64_t* __restrict__ yoff, const double* __restrict__ Z,
                                                   int kp, double alpha, const int* __restrict__ gate) {
  const int d = blockIdx.y;
  if (gate && *gate == 0) return;
  const int Ra = (mrows[d] + 63) / 64, Ca = (mcols[d] + 63) / 64;
  const int Ro = (mode != 0 && op == 1) ? Ca : Ra;
  const int Ko = (mode != 0 && op == 1) ? Ra : Ca;
  const int i = blockIdx.x;
  if (i >= Ro) return;
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
  const double* Xd = X + xoff[d] + c0;
  for (int j = 0; j < Ko; j++) {
    bool tr = false;
    const int64_t ti = mode == 2 ? 0 : tile_of(mode, op, Ra, i, j, &tr);
    const double* tile = Ad + ti * 4096;
    for (int k0 = 0; k0 < 64; k0 += 32) {
      if (mode == 2)
        load_chunk_cm<double>(Ad, lda[d], mrows[d], mcols[d], op, i, j, k0, As, Xd + (int64_t)(j * 64) * kp, kp, Xs);
      else
        load_chunk<double>(tile, tr, k0, As, Xd + (int64_t)(j * 64) * kp, kp, Xs);
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < 32; kk += 4) {
        const double a = As[(kk + (lane >> 4)) * 65 + rb + (lane & 15)];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const double b = Xs[(kk + (lane >> 4)) * 64 + 16 * c + (lane & 15)];
          acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
        }
      }
      __syncthreads();
    }
  }
  double* Yd = Y + yoff[d] + c0;
  const double* Zd = Z ? Z + yoff[d] + c0 : nullptr;
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = i * 64 + rb + 4 * r + (lane >> 4);
      const int col = 16 * c + (lane & 15);
      double v = alpha * acc[c][r];
      if (Zd) v += Zd[(int64_t)row * kp + col];
      Yd[(int64_t)row * kp + col] = v;
    }
}

hipError_t launch_tgemm_f32(hipStream_t st, int nsub, int maxRo, int mode, int op, const int* mrows, const int* mcols,
                            const float* A, const int64_t* aoff, const int* lda, const float* X, const int64_t* xoff, float* Y,
                            const int64_t* yoff, const float* Z, int kp, float alpha, const int* gate) {
  if (kp % 64) return hipErrorInvalidValue;
  if (mode == 2 && !lda) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tgemm_f32, dim3(maxRo, nsub, kp / 64), dim3(256), 0, st, mode, op, mrows, mcols, A, aoff, lda, X,
                     xoff, Y, yoff, Z, kp, alpha, gate);
  return hipGetLastError();
}

hipError_t launch_tgemm_f64(hipStream_t st, int nsub, int maxRo, int mode, int op, const int* mrows, const int* mcols,
                            const double* A, const int64_t* aoff, const int* lda, const double* X, const int64_t* xoff, double* Y,
                            const int64_t* yoff, const double* Z, int kp, double alpha, const int* gate) {
  if (kp % 64) return hipErrorInvalidValue;
  if (mode == 2 && !lda) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_tgemm_f64, dim3(maxRo, nsub, kp / 64), dim3(256), 0, st, mode, op, mrows, mcols, A, aoff, lda, X,
                     xoff, Y, yoff, Z, kp, alpha, gate);
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

}  // namespace hfls
