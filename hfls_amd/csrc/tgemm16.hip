// Split-FP16 tile GEMM of the mixed-precision block solve: Y_d = alpha A_d X_d
// for the symmetric per-subdomain operators (deflated K+_bb, S_i), packed lower
// 64x64 tiles as in tgemm.hip (mode 0, op 0), FP32 result.
//
// Each FP32 value v is held as two FP16 values, hi = fp16(v/s) and
// lo = fp16(v/s - hi), with a power-of-two scale s: s_A per subdomain for the
// matrix (a power of two that maps max |A_d| into [2^14, 2^15)), s_X per
// subdomain and column for the right-hand sides (the same rule over the column).
// hi + lo carries 22 significand bits; the product takes the three terms
// hi*hi + hi*lo + lo*hi (the lo*lo term is below 2^-22 relative) on the FP16
// matrix cores (v_mfma_f32_32x32x16_f16, FP32 accumulation), so one K chunk costs
// 6 matrix-core operations of 32x32x16 instead of 16 of 32x32x2 in FP32, with the
// same HBM bytes per matrix element (two FP16 planes in the 16 KB of an FP32 tile).
//
// Tile storage after kb_split16: each 64x64 tile keeps its 16 KB; halves
// 0..4095 are the hi plane and 4096..8191 the lo plane, each column major inside
// the tile like the FP32 tile it replaces.
//
// Operand lane layout of v_mfma_f32_32x32x16_f16: lane l supplies 8 values of
// A row l%32 and of B column l%32, k group l/32.  A and B take the same
// k = kk + 8*(l/32) + s for slot s, so the 16 products of a row and a column are
// summed whatever order the hardware gives the slots; the result layout is that
// of the FP32 32x32 operation (row 8(r/4)+4(l/32)+r%4, column l%32).
#include "kernels.hpp"

namespace hfls {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// power of two s with m / s in [2^14, 2^15) (1 for m == 0), at least 2^-126 so that
// 1 / s stays finite (values below 2^-111 then keep fewer significant bits)
static __device__ __forceinline__ float pow2_scale(float m) {
  if (!(m > 0.0f)) return 1.0f;
  int e;
  (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
  return ldexpf(1.0f, max(e - 15, -126));
}

static __device__ __forceinline__ float block_max256(float m, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// scale[d] from the FP32 tiles of subdomain d (T (T + 1) / 2 tiles from poff[d],
// T = ceil(ns[d] / 64))
__global__ __launch_bounds__(256) void kb_tile_scale(const int* __restrict__ ns, const int64_t* __restrict__ poff,
                                                     const float* __restrict__ P32, float* __restrict__ scale) {
  __shared__ float red[4];
  const int d = blockIdx.x;
  const int64_t T = (ns[d] + 63) / 64;
  const int64_t beg = poff[d], end = beg + T * (T + 1) / 2 * 4096;
  float m = 0.0f;
  for (int64_t e = beg + 4 * threadIdx.x; e < end; e += 1024) {
    const f4 v = *(const f4*)(P32 + e);
    m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
  }
  m = block_max256(m, red);
  if (threadIdx.x == 0) scale[d] = pow2_scale(m);
}

// in place: FP32 tile -> hi and lo FP16 planes of v / scale[d]
__global__ __launch_bounds__(256) void kb_split16(const int* __restrict__ ns, const int64_t* __restrict__ poff,
                                                  const float* __restrict__ scale, float* __restrict__ P32) {
  const int d = blockIdx.y;
  const int T = (ns[d] + 63) / 64;
  if ((int)blockIdx.x >= T * (T + 1) / 2) return;
  float* tile = P32 + poff[d] + (int64_t)blockIdx.x * 4096;
  const float inv = 1.0f / scale[d];
  f4 v[4];
#pragma unroll
  for (int u = 0; u < 4; u++) v[u] = *(const f4*)(tile + 4 * threadIdx.x + 1024 * u);
  __syncthreads();
  _Float16* h = (_Float16*)tile;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    half4 hi, lo;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float x = v[u][q] * inv;
      hi[q] = (_Float16)x;
      lo[q] = (_Float16)(x - (float)hi[q]);
    }
    const int e = 4 * threadIdx.x + 1024 * u;
    *(half4*)(h + e) = hi;
    *(half4*)(h + 4096 + e) = lo;
  }
}

// sx[d*kp + c] = scale of column c of X_d (rows 0 .. ns[d]-1); 16 waves per
// subdomain, each lane with 8 independent loads in flight per step
__global__ __launch_bounds__(1024) void kb_colscale(const int* __restrict__ ns, const int64_t* __restrict__ xoff,
                                                    int kp, const float* __restrict__ X, float* __restrict__ sx,
                                                    const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ float red[16][64];
  const int d = blockIdx.x, c = blockIdx.y * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
  const float* Xd = X + xoff[d] + c;
  const int n = ns[d];
  float m[8];
#pragma unroll
  for (int u = 0; u < 8; u++) m[u] = 0.0f;
  int r = g;
  for (; r + 112 < n; r += 128) {  // rows r + 16 u, u = 0..7, all below n
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = Xd[(int64_t)(r + 16 * u) * kp];
#pragma unroll
    for (int u = 0; u < 8; u++) m[u] = fmaxf(m[u], fabsf(v[u]));
  }
  for (; r < n; r += 16) m[0] = fmaxf(m[0], fabsf(Xd[(int64_t)r * kp]));
#pragma unroll
  for (int u = 1; u < 8; u++) m[0] = fmaxf(m[0], m[u]);
  red[g][threadIdx.x & 63] = m[0];
  __syncthreads();
  if (g == 0) {
    float v = red[0][threadIdx.x];
#pragma unroll
    for (int q = 1; q < 16; q++) v = fmaxf(v, red[q][threadIdx.x]);
    sx[(int64_t)d * kp + c] = pow2_scale(v);
  }
}

// One workgroup: 64 output rows (row block i) x 64 columns, K in chunks of KC (32:
// half a tile, 64: a whole tile) through LDS, A and X split into hi/lo as they are
// staged; the workgroup order and the rotated tile order are those of k_tgemm_f32
// (all row blocks of a subdomain on one XCD, so X_d is fetched once per launch; tile
// columns from (i & ~1)).  Every stored off-diagonal tile is read by two row blocks;
// FETCH_SIZE shows both reads come from HBM (2 x the packed tiles), so the kernel is
// bound by that traffic.  Both chunk sizes issue the same matrix-core operations in
// the same order, so they give bit-identical results.
template <int KC, int PF>
__global__ __launch_bounds__(256) void k_tgemm_h16(int nsub, int nslot, const int* __restrict__ order, int norder,
                                                   const int* __restrict__ mrows, const float* __restrict__ A,
                                                   const int64_t* __restrict__ aoff, const float* __restrict__ sA,
                                                   const float* __restrict__ X, const int64_t* __restrict__ xoff,
                                                   const float* __restrict__ sx, int kp, float alpha,
                                                   const int* __restrict__ gate, float* __restrict__ Yf, int exact) {
  constexpr int LDH = KC + 8;  // LDS row: KC halves + 8 of padding
  constexpr int NQ = KC / 32;  // 32-deep slices per chunk
  constexpr int CPT = 64 / KC; // chunks per tile
  const int sl = blockIdx.x >> 3;
  const int gi = (sl / nslot) * 8 + (blockIdx.x & 7);
  const int d = order ? (gi < norder ? order[gi] : -1) : (gi < nsub ? gi : -1);
  if (d < 0) return;
  if (gate && *gate == 0) return;
  const int T = (mrows[d] + 63) / 64;
  const int i0 = sl % nslot;
  if (i0 >= T) return;
  // K chunk c covers tile column (c / CPT - rot) mod T; exact > 0: odd row blocks shifted by exact - 1 steps
  const int rot = (exact ? i0 + (exact - 1) * (i0 & 1) : (i0 & ~1)) % T;
  const int c0 = blockIdx.z * 64;
  __shared__ __attribute__((aligned(16))) _Float16 Ah[64 * LDH], Al[64 * LDH], Xh[64 * LDH], Xl[64 * LDH];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int rb = (w >> 1) * 32, cb = (w & 1) * 32;
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  const _Float16* Ad = (const _Float16*)(A + aoff[d]);
  const float* Xd = X + xoff[d] + c0;
  // X staging: column xc, rows k0 + 32 q + 8 xg .. + 7 of the chunk
  const int xc = t & 63, xg = t >> 6;
  const float xinv = 1.0f / sx[(int64_t)d * kp + c0 + xc];
  // one staged K chunk in registers: A tile slices (hi, lo), X rows, orientation
  struct Stage {
    half8 sah[NQ], sal[NQ];
    float sxv[NQ][8];
    bool tr;
  };
  auto fetch = [&](int c, Stage& g) {
    int tj = c / CPT - rot;
    if (tj < 0) tj += T;
    const int k0 = (c % CPT) * KC;
    g.tr = i0 < tj;  // op(A)(i0, tj) is stored tile (tj, i0) transposed
    const int64_t ti = g.tr ? (int64_t)i0 * T - (int64_t)i0 * (i0 - 1) / 2 + (tj - i0)
                            : (int64_t)tj * T - (int64_t)tj * (tj - 1) / 2 + (i0 - tj);
    const _Float16* tile = Ad + ti * 8192;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      // direct: 8 rows (t&7)*8.. of column k0 + 32q + (t>>3); transposed: 8 k of row t>>2
      const int k = k0 + 32 * q;
      const int64_t e = g.tr ? (int64_t)(t >> 2) * 64 + k + (t & 3) * 8 : (int64_t)(k + (t >> 3)) * 64 + (t & 7) * 8;
      g.sah[q] = *(const half8*)(tile + e);
      g.sal[q] = *(const half8*)(tile + 4096 + e);
      const float* xr = Xd + (int64_t)(tj * 64 + k + xg * 8) * kp + xc;
#pragma unroll
      for (int u = 0; u < 8; u++) g.sxv[q][u] = xr[(int64_t)u * kp];
    }
  };
  auto stash = [&](const Stage& g) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      if (g.tr) {
        const int o = (t >> 2) * LDH + 32 * q + (t & 3) * 8;
        *(half8*)(Ah + o) = g.sah[q];
        *(half8*)(Al + o) = g.sal[q];
      } else {
        const int kk = 32 * q + (t >> 3), r8 = (t & 7) * 8;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          Ah[(r8 + u) * LDH + kk] = g.sah[q][u];
          Al[(r8 + u) * LDH + kk] = g.sal[q][u];
        }
      }
      half8 h, l;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const float x = g.sxv[q][u] * xinv;
        h[u] = (_Float16)x;
        l[u] = (_Float16)(x - (float)h[u]);
      }
      *(half8*)(Xh + xc * LDH + 32 * q + xg * 8) = h;
      *(half8*)(Xl + xc * LDH + 32 * q + xg * 8) = l;
    }
  };
  const int ao = (rb + (lane & 31)) * LDH + 8 * (lane >> 5);
  const int bo = (cb + (lane & 31)) * LDH + 8 * (lane >> 5);
  auto compute = [&]() {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 16) {
      const half8 ah = *(const half8*)(Ah + ao + kk), al = *(const half8*)(Al + ao + kk);
      const half8 bh = *(const half8*)(Xh + bo + kk), bl = *(const half8*)(Xl + bo + kk);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc, 0, 0, 0);
    }
  };
  const int nch = CPT * T;
  Stage s0, s1;
  fetch(0, s0);
  if (PF == 2 && nch > 1) fetch(1, s1);
  stash(s0);
  __syncthreads();
  if (PF == 1) {
    // the next chunk's loads are in flight while the current one is multiplied
    for (int c = 0; c < nch; c++) {
      const bool more = c + 1 < nch;
      if (more) fetch(c + 1, s0);
      compute();
      __syncthreads();
      if (more) stash(s0);
      __syncthreads();
    }
  } else {
    // the loads of the two chunks after the current one are in flight; the loop
    // body covers two chunks so that both register stages keep static names
    for (int c = 0; c < nch; c += 2) {
      if (c + 2 < nch) fetch(c + 2, s0);  // s0 held chunk c, already in LDS
      compute();
      __syncthreads();
      if (c + 1 >= nch) break;
      stash(s1);  // chunk c + 1
      __syncthreads();
      if (c + 3 < nch) fetch(c + 3, s1);
      compute();
      __syncthreads();
      if (c + 2 < nch) {
        stash(s0);  // chunk c + 2
        __syncthreads();
      }
    }
  }
  const int col = cb + (lane & 31);
  // the two scales applied one after the other: their product alone may underflow
  const float fa = alpha * sA[d], fx = sx[(int64_t)d * kp + c0 + col];
  float* Yd = Yf + xoff[d] + c0;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int row = i0 * 64 + rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    Yd[(int64_t)row * kp + col] = (acc[r] * fa) * fx;
  }
}

hipError_t launch_split16(hipStream_t st, int nsub, int maxtiles, const int* ns, const int64_t* poff, float* P32,
                          float* scale) {
  if (nsub <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_tile_scale, dim3(nsub), dim3(256), 0, st, ns, poff, (const float*)P32, scale);
  hipLaunchKernelGGL(kb_split16, dim3(maxtiles, nsub), dim3(256), 0, st, ns, poff, (const float*)scale, P32);
  return hipGetLastError();
}

hipError_t launch_tgemm_h16(hipStream_t st, int nsub, int maxT, const int* ns, const float* A, const int64_t* aoff,
                            const float* sA, const float* X, const int64_t* xoff, float* sx, int kp, float alpha,
                            const int* gate, float* Yf, const int* xcd_order, int norder, int kc, int scales,
                            int lds_pad, int exact, int pf) {
  if (lds_pad < 0 || lds_pad > 120 * 1024 || (pf != 1 && pf != 2)) return hipErrorInvalidValue;
  if (nsub <= 0 || maxT <= 0) return hipSuccess;
  if (kp <= 0 || kp % 64 || (kc != 32 && kc != 64)) return hipErrorInvalidValue;
  if (xcd_order && (norder < 8 || norder % 8)) return hipErrorInvalidValue;
  if (scales) hipLaunchKernelGGL(kb_colscale, dim3(nsub, kp / 64), dim3(1024), 0, st, ns, xoff, kp, X, sx, gate);
  const int groups = xcd_order ? norder / 8 : (nsub + 7) / 8;  // subdomain groups of 8, one per XCD
  // lds_pad: unused dynamic LDS per workgroup, which caps the workgroups resident per CU;
  // pf: K chunks whose loads are in flight while one is multiplied (1 or 2)
  const dim3 grid(8 * maxT * groups, 1, kp / 64);
#define HFLS_TG16(KCv, PFv)                                                                                  \
  hipLaunchKernelGGL((k_tgemm_h16<KCv, PFv>), grid, dim3(256), lds_pad, st, nsub, maxT, xcd_order, norder, ns, A, \
                     aoff, sA, X, xoff, (const float*)sx, kp, alpha, gate, Yf, exact)
  if (kc == 64 && pf == 2) HFLS_TG16(64, 2);
  else if (kc == 64) HFLS_TG16(64, 1);
  else if (pf == 2) HFLS_TG16(32, 2);
  else HFLS_TG16(32, 1);
#undef HFLS_TG16
  return hipGetLastError();
}

// the per-column scales of X for the band form of the GEMM (tgemm16b.hip)
hipError_t launch_kb_colscale(hipStream_t st, int nsub, const int* ns, const int64_t* xoff, int kp, const float* X,
                              float* sx, const int* gate) {
  if (nsub <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_colscale, dim3(nsub, kp / 64), dim3(1024), 0, st, ns, xoff, kp, X, sx, gate);
  return hipGetLastError();
}

}  // namespace hfls
