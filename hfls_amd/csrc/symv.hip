// Symmetric dense operators (K+_bb, S_i) stored as packed lower 64x64 tiles.
//
// Per subdomain, T = ceil(n_s/64) tile rows; tile (ti,tj), ti >= tj, is stored at
// tile index tj*T - tj*(tj-1)/2 + (ti-tj), column major inside the tile, zero
// padded.  Only the lower triangle of tiles is streamed from HBM: an off-diagonal
// tile contributes T x_tj to y_ti and T^T x_ti to y_tj, halving the bytes of the
// dense F_i / S_i stream of every PCPG iteration.
//
// k_symv_tiles: one workgroup per (subdomain, tile column tj), longest columns
//   first.  Thread t holds a 2-row x 8-column piece of each tile (rows 2(t%32), +1; columns t/32 + 8k), read
//   with one 16-byte load per column (coalesced 4 KB per instruction across the
//   workgroup); the next tile is loaded before the current one is reduced.
//   Row parts are reduced across the 8 column groups through LDS and written as
//   partials; column parts accumulate in registers over the whole tile column.
// k_symv_reduce: y[ti*64 + r] = sum_{tj <= ti} partial(tj, ti)[r] in fixed order.
#include "kernels.hpp"

namespace hfls {

static __device__ __forceinline__ int64_t tile_index(int T, int tj, int ti) {
  return (int64_t)tj * T - (int64_t)tj * (tj - 1) / 2 + (ti - tj);
}

// pack a compact symmetric matrix (ld) into lower tiles
__global__ __launch_bounds__(256) void k_pack_tiles(const int* __restrict__ ns, const double* __restrict__ A,
                                                    const int64_t* __restrict__ aoff, const int* __restrict__ lda,
                                                    double* __restrict__ P, const int64_t* __restrict__ poff) {
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
  const double* Ad = A + aoff[d];
  const int64_t l = lda[d];
  double* tile = P + poff[d] + tile_index(T, tj, ti) * 4096;
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int c = e >> 6, r = e & 63;
    const int gr = 64 * ti + r, gc = 64 * tj + c;
    tile[e] = (gr < n && gc < n) ? Ad[gr + (int64_t)gc * l] : 0.0;
  }
}

hipError_t launch_pack_tiles(hipStream_t st, int nsub, int maxtiles, const int* ns, const double* A,
                             const int64_t* aoff, const int* lda, double* P, const int64_t* poff) {
  hipLaunchKernelGGL(k_pack_tiles, dim3(maxtiles, nsub), dim3(256), 0, st, ns, A, aoff, lda, P, poff);
  return hipGetLastError();
}

// x_b = B^T (wgt .* v) per packed b position (the interface gather)
__global__ __launch_bounds__(256) void k_gather(int64_t nb, const int* __restrict__ gptr, const int* __restrict__ gidx,
                                                const double* __restrict__ v, const double* __restrict__ wgt,
                                                double* __restrict__ xb, const int* __restrict__ done) {
  if (done && *done) return;
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= nb) return;
  double s = 0.0;
  const int e1 = gptr[q + 1];
  for (int e = gptr[q]; e < e1; e++) {
    const int code = gidx[e];
    const int r = code >= 0 ? code : ~code;
    double val = v[r];
    if (wgt) val *= wgt[r];
    s += code >= 0 ? val : -val;
  }
  xb[q] = s;
}

hipError_t launch_gather(hipStream_t st, int64_t nb, const int* gptr, const int* gidx, const double* v,
                         const double* wgt, double* xb, const int* done) {
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, nb, gptr, gidx, v, wgt, xb, done);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_symv_tiles(const int* __restrict__ ns, const double* __restrict__ P,
                                                    const int64_t* __restrict__ poff, const double* __restrict__ x,
                                                    const int64_t* __restrict__ xoff, double* __restrict__ part,
                                                    const int64_t* __restrict__ partoff,
                                                    const int* __restrict__ done) {
  if (done && *done) return;
  const int d = blockIdx.x;
  const int n = ns[d];
  const int T = (n + 63) / 64;
  const int tj = blockIdx.y;
  if (tj >= T) return;
  __shared__ double red[8 * 64];
  __shared__ double xrow[64];
  const int t = threadIdx.x;
  const int cg = t >> 5;         // column group 0..7: columns cg + 8k
  const int rp = t & 31;         // row pair: rows 2rp, 2rp+1
  const double* xd = x + xoff[d];
  double xc[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int gc = 64 * tj + cg + 8 * k;
    xc[k] = gc < n ? xd[gc] : 0.0;
  }
  double colacc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double diagrow = 0.0;
  const double* Pd = P + poff[d];
  double* pd = part + partoff[d];
  const double* tile = Pd + tile_index(T, tj, tj) * 4096;
  double2 cur[8], nxt[8];
#pragma unroll
  for (int k = 0; k < 8; k++) cur[k] = *reinterpret_cast<const double2*>(tile + 2 * t + 512 * k);
  for (int ti = tj; ti < T; ti++) {
    if (ti + 1 < T) {
      const double* tn = tile + 4096;  // tiles of one tile column are contiguous
#pragma unroll
      for (int k = 0; k < 8; k++) nxt[k] = *reinterpret_cast<const double2*>(tn + 2 * t + 512 * k);
    }
    const int r0 = 64 * ti + 2 * rp;
    double p0 = 0.0, p1 = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      p0 += cur[k].x * xc[k];
      p1 += cur[k].y * xc[k];
    }
    if (ti > tj) {
      const double x0 = r0 < n ? xd[r0] : 0.0;
      const double x1 = r0 + 1 < n ? xd[r0 + 1] : 0.0;
#pragma unroll
      for (int k = 0; k < 8; k++) colacc[k] += cur[k].x * x0 + cur[k].y * x1;
    }
    red[cg * 64 + 2 * rp] = p0;
    red[cg * 64 + 2 * rp + 1] = p1;
    __syncthreads();
    if (t < 64) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 8; k++) s += red[k * 64 + t];
      if (ti > tj) pd[tile_index(T, tj, ti) * 64 + t] = s;
      else diagrow = s;
    }
    __syncthreads();
    tile += 4096;
#pragma unroll
    for (int k = 0; k < 8; k++) cur[k] = nxt[k];
  }
  // column sums over the 32 row pairs of each column group (lanes 0..31 / 32..63 of a wave)
#pragma unroll
  for (int k = 0; k < 8; k++) {
    double v = colacc[k];
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    colacc[k] = v;
  }
  if (rp == 0) {
#pragma unroll
    for (int k = 0; k < 8; k++) xrow[cg + 8 * k] = colacc[k];
  }
  __syncthreads();
  if (t < 64) pd[tile_index(T, tj, tj) * 64 + t] = diagrow + xrow[t];
}

hipError_t launch_symv_tiles(hipStream_t st, int nsub, int maxT, const int* ns, const double* P, const int64_t* poff,
                             const double* x, const int64_t* xoff, double* part, const int64_t* partoff,
                             const int* done) {
  // subdomains fastest: every subdomain's longest tile column (tj = 0) is dispatched
  // first and the one-tile columns last, and with the hardware handing consecutive
  // workgroups to the 8 XCDs in turn, all columns of subdomain d run on XCD d mod 8
  // (nsub a multiple of 8), sharing one L2 for x
  hipLaunchKernelGGL(k_symv_tiles, dim3(nsub, maxT), dim3(256), 0, st, ns, P, poff, x, xoff, part, partoff, done);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_symv_reduce(int64_t nb, const int* __restrict__ bsub,
                                                     const int* __restrict__ ns, const int64_t* __restrict__ boff,
                                                     const double* __restrict__ part,
                                                     const int64_t* __restrict__ partoff, double* __restrict__ y,
                                                     const int* __restrict__ done) {
  if (done && *done) return;
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= nb) return;
  const int d = bsub[q];
  const int p = (int)(q - boff[d]);
  const int T = (ns[d] + 63) / 64;
  const int ti = p >> 6, r = p & 63;
  const double* pd = part + partoff[d];
  double s = 0.0;
  for (int tj = 0; tj <= ti; tj++) s += pd[tile_index(T, tj, ti) * 64 + r];
  y[q] = s;
}

hipError_t launch_symv_reduce(hipStream_t st, int64_t nb, const int* bsub, const int* ns, const int64_t* boff,
                              const double* part, const int64_t* partoff, double* y, const int* done) {
  hipLaunchKernelGGL(k_symv_reduce, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, nb, bsub, ns, boff, part,
                     partoff, y, done);
  return hipGetLastError();
}

}  // namespace hfls
