// Split-FP16 tile GEMM, band form: Y_d = alpha A_d X_d for the symmetric packed
// tiles of tgemm16.hip (deflated K+_bb, S_i; hi and lo FP16 planes per 16 KB tile,
// power-of-two scales per subdomain and per column of X) with every stored tile
// read from HBM once.
//
// The row form (k_tgemm_h16) gives each 64-row output block its own workgroup,
// so every off-diagonal tile T(i,j) is streamed twice, once for T X_j into row
// block i and once for T^T X_i into row block j, and it runs at the fabric rate of
// that doubled traffic.  Here one 256-thread workgroup owns a band of M row blocks
// (i0 .. i0+nb-1) of one subdomain and visits every stored tile T(i,j) of those
// rows (j <= i) once; each of its 4 waves computes one 32x32 quadrant of both
//   * T(i,j) X_j, added to the band's row block i, and
//   * T(i,j)^T X_i (i > j), added to row block j: inside the band directly, for
//     j < i0 into a partial of this band, written out when column j is done,
// with 3 FP16 products (hi*hi, hi*lo, lo*hi) per 16-deep K step.  Row block j then
// receives the partials of the bands below it in k_tgemm_h16b_reduce, added in
// ascending band order, so the result is bitwise reproducible.  Partial traffic:
// about 2/M of the tile bytes (written once, read once).
//
// The tile is staged once, in its stored column-major layout: the T^T X_i operand
// (8 consecutive rows of a tile column per lane) is a 16-byte load, the T X_j
// operand (8 consecutive columns of a tile row) two transposing loads
// (ds_read_b64_tr_b16: in each group of 16 lanes, lane l receives column l of the
// 4 x 16 block of 16-bit values whose rows the 16 lanes address).  Tile columns are
// XOR-swizzled in groups of 8 rows (see tpos) so that the staging stores, the
// 16-byte loads and the transposing loads are all free of bank conflicts.  X_i is
// staged per step; X_j of the next off-band column is staged a quarter per step
// during the current one (two column buffers).  70 KB of LDS: two workgroups per
// CU.
//
// Bands are aligned to the last tile row (band b covers rows T-(B-b)M ..
// T-(B-1-b)M-1, band 0 the rest), so every band with off-band columns has exactly
// M rows and the step sequence is static: the off-band part is a loop over columns
// j < i0 with the M rows unrolled, the in-band triangle is unrolled; accumulator
// indices are compile-time and no load sits behind a branch.
//
// Partial slots of subdomain d (from poff[d], 4096 floats per slot and 64-column
// block of X, in the accumulator order ((wave * 16 + r) * 64 + lane)): band b >= 1
// owns slots pbase(b) + j for j < i0(b), pbase(b) = (b-1) nb0 + M (b-1)(b-2)/2,
// nb0 = T - (B-1) M the rows of band 0.
#include <type_traits>

#include "kernels.hpp"

namespace hfls {

// f(integral_constant<int, I>) for I = B .. E-1, every index a compile-time constant
// (accumulator arrays indexed by them stay in registers)
template <int B, int E, class F>
static __device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// step e of a triangle enumerated column-major over (jj, ii), jj <= ii < M
template <int M>
static constexpr int tri_jj(int e) {
  int jj = 0;
  while (e >= M - jj) {
    e -= M - jj;
    jj++;
  }
  return jj;
}
template <int M>
static constexpr int tri_ii(int e) {
  const int jj = tri_jj<M>(e);
  int base = 0;
  for (int q = 0; q < jj; q++) base += M - q;
  return jj + (e - base);
}

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((__vector_size__(4 * sizeof(__fp16)))) __fp16 fp16x4;

constexpr int BM = HFLS_H16B_M;  // row blocks per band

// position of tile element (row r, column c) in an LDS tile plane: column major,
// rows XORed in groups of 8 by f(c) = 16 (c mod 4) ^ 8 ((c / 4) mod 2), so that the
// 4 columns one transposing load spans, and the 8 columns 8 lanes of a 16-byte
// load span, fall in different bank groups
static __device__ __forceinline__ int tpos(int c, int r) { return c * 64 + (r ^ (((c & 3) << 4) ^ (((c >> 2) & 1) << 3))); }

static __device__ __forceinline__ half8 join8(fp16x4 a, fp16x4 b) {
  half8 o;
  o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = a[3];
  o[4] = b[0]; o[5] = b[1]; o[6] = b[2]; o[7] = b[3];
  return o;
}

struct BandGeom {
  int T, B, nb0;
};
static __host__ __device__ __forceinline__ BandGeom band_geom(int ns) {
  BandGeom g;
  g.T = (ns + 63) / 64;
  g.B = (g.T + BM - 1) / BM;
  g.nb0 = g.T - (g.B - 1) * BM;
  return g;
}
// first partial slot of band b >= 1 (one slot per column j < i0(b))
static __host__ __device__ __forceinline__ int64_t band_pbase(const BandGeom& g, int b) {
  return (int64_t)(b - 1) * g.nb0 + (int64_t)BM * (b - 1) * (b - 2) / 2;
}

template <int M>
__global__ __launch_bounds__(256, 2) void k_tgemm_h16b(int nsub, int nslot, const int* __restrict__ order, int norder,
                                                    const int* __restrict__ mrows, const float* __restrict__ A,
                                                    const int64_t* __restrict__ aoff, const float* __restrict__ sA,
                                                    const float* __restrict__ X, const int64_t* __restrict__ xoff,
                                                    const float* __restrict__ sx, int kp, float alpha,
                                                    const int* __restrict__ gate, float* __restrict__ Yf,
                                                    float* __restrict__ part, const int64_t* __restrict__ poff) {
  static_assert(M == 4, "next-column staging: 4 steps x 4 waves x 4 rows");
  constexpr int LDH = 72;       // padded row of an X plane (halves)
  constexpr int XB = 64 * LDH;  // one X plane
  constexpr int TP = 4096;      // one tile plane (swizzled)
  __shared__ __attribute__((aligned(16))) _Float16 lds[2 * TP + 6 * XB];
  _Float16* const Sh = lds;            // tile, hi
  _Float16* const Sl = lds + TP;       //       lo
  _Float16* const Xih = lds + 2 * TP;  // X_i of the step, hi; lo = +XB
  auto xj_ptr = [&](int col) { return lds + 2 * TP + 2 * XB + (col & 1) * 2 * XB; };  // X_j, hi; lo = +XB

  const int sl = blockIdx.x >> 3;
  const int gi = (sl / nslot) * 8 + (blockIdx.x & 7);
  const int d = order ? (gi < norder ? order[gi] : -1) : (gi < nsub ? gi : -1);
  if (d < 0) return;
  if (gate && *gate == 0) return;
  const BandGeom geo = band_geom(mrows[d]);
  const int T = geo.T;
  const int bb = sl % nslot;
  if (bb >= geo.B) return;
  const int b = geo.B - 1 - bb;  // the heaviest bands of a subdomain start first
  const int i0 = b == 0 ? 0 : geo.nb0 + (b - 1) * M;
  const int nb = b == 0 ? geo.nb0 : M;
  const int c0 = blockIdx.z * 64;

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int rb = (w >> 1) * 32, cb = (w & 1) * 32;
  floatx16 accr[M], accc[M];  // row block i0 + m: T X_j sums; column i0 + m (accc[0]: the off-band partial)
  static_for<0, M>([&](auto mc) {
#pragma unroll
    for (int r = 0; r < 16; r++) accr[mc.value][r] = accc[mc.value][r] = 0.0f;
  });

  const _Float16* Ad = (const _Float16*)(A + aoff[d]);
  const float* Xd = X + xoff[d] + c0;
  // tile staging: 16-byte units t and t + 256 of each plane, column (t >> 3) (+32),
  // rows (t & 7) * 8 .. + 7
  const int tc = t >> 3, tr8 = (t & 7) * 8;
  const int sp0 = tpos(tc, tr8), sp1 = tpos(tc + 32, tr8);
  const _Float16* const tsrc = Ad + tc * 64 + tr8;
  // X staging: column xn, rows xk .. xk + 15
  const int xn = t & 63, xk = w * 16;
  const float xinv = 1.0f / sx[(int64_t)d * kp + c0 + xn];
  const float* const xsrc = Xd + (int64_t)xk * kp + xn;

  struct Stage {
    half8 h0, h1, l0, l1;  // tile
    float xi[16];          // X_i rows xk .. xk + 15
    float xn4[4];          // 4 values of the next off-band column's X block
  };
  // loads of step (row block i, column j) and X rows xrow .. xrow + 3 (column xn)
  auto fetch = [&](Stage& g, int i, int j, int xrow) {
    const int64_t ti = (int64_t)j * T - (int64_t)j * (j - 1) / 2 + (i - j);
    const _Float16* tile = tsrc + ti * 8192;
    g.h0 = *(const half8*)tile;
    g.h1 = *(const half8*)(tile + 2048);
    g.l0 = *(const half8*)(tile + 4096);
    g.l1 = *(const half8*)(tile + 6144);
    const float* xr = xsrc + (int64_t)i * 64 * kp;
#pragma unroll
    for (int u = 0; u < 16; u++) g.xi[u] = xr[(int64_t)u * kp];
#pragma unroll
    for (int u = 0; u < 4; u++) g.xn4[u] = Xd[(int64_t)(xrow + u) * kp + xn];
  };
  auto split8 = [&](const float* v, _Float16* h, int k) {  // 8 values -> rows k .. k+7 of column xn
    half8 hh, ll;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const float x = v[u] * xinv;
      hh[u] = (_Float16)x;
      ll[u] = (_Float16)(x - (float)hh[u]);
    }
    *(half8*)(h + xn * LDH + k) = hh;
    *(half8*)(h + XB + xn * LDH + k) = ll;
  };
  auto stash = [&](const Stage& g) {
    *(half8*)(Sh + sp0) = g.h0;
    *(half8*)(Sh + sp1) = g.h1;
    *(half8*)(Sl + sp0) = g.l0;
    *(half8*)(Sl + sp1) = g.l1;
    split8(g.xi, Xih, xk);
    split8(g.xi + 8, Xih, xk + 8);
  };
  auto stash_next = [&](const Stage& g, int col, int k) {  // rows k .. k+3 of column col's X_j
    _Float16* h = xj_ptr(col);
    half4 hh, ll;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const float x = g.xn4[u] * xinv;
      hh[u] = (_Float16)x;
      ll[u] = (_Float16)(x - (float)hh[u]);
    }
    *(half4*)(h + xn * LDH + k) = hh;
    *(half4*)(h + XB + xn * LDH + k) = ll;
  };
  // operand offsets.  T X_j: lane group g4 = lane/16 reads rows m0 + li of columns
  // kg .. kg+7 as two transposing loads (kg = kk + 8 (g4/2), m0 = rb + 16 (g4%2));
  // lane li addresses column kg + li/4 (+4), rows m0 + 4 (li%4) .. +3.
  const int g4 = lane >> 4, li = lane & 15;
  const int trow = rb + 16 * (g4 & 1) + 4 * (li & 3), tcol = 8 * (g4 >> 1) + (li >> 2);
  const int tro1 = tpos(tcol, trow), tro2 = tpos(tcol + 4, trow);  // + 64 kk
  // T^T X_i: lane reads tile column rb + lane%32, rows kk + 8 (lane/32) .. +7
  const int ccol = rb + (lane & 31), ck = 8 * (lane >> 5);
  const int bo = (cb + (lane & 31)) * LDH + ck;  // X operand: column cb + lane%32, rows kk + 8 (lane/32)
  auto tr4 = [](const _Float16* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4*)p);
  };
  auto mm = [&](floatx16& ar, const _Float16* xj, floatx16* ac, const bool col) {
#pragma unroll
    for (int kk = 0; kk < 64; kk += 16) {
      const fp16x4 r1h = tr4(Sh + tro1 + 64 * kk), r2h = tr4(Sh + tro2 + 64 * kk);
      const fp16x4 r1l = tr4(Sl + tro1 + 64 * kk), r2l = tr4(Sl + tro2 + 64 * kk);
      const half8 ah = join8(r1h, r2h), al = join8(r1l, r2l);
      const half8 bh = *(const half8*)(xj + bo + kk), bl = *(const half8*)(xj + XB + bo + kk);
      ar = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, ar, 0, 0, 0);
      ar = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, ar, 0, 0, 0);
      ar = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, ar, 0, 0, 0);
      if (col) {
        const int o = tpos(ccol, kk + ck);
        const half8 ch = *(const half8*)(Sh + o), cl = *(const half8*)(Sl + o);
        const half8 xh = *(const half8*)(Xih + bo + kk), xl = *(const half8*)(Xih + XB + bo + kk);
        *ac = __builtin_amdgcn_mfma_f32_32x32x16_f16(ch, xh, *ac, 0, 0, 0);
        *ac = __builtin_amdgcn_mfma_f32_32x32x16_f16(ch, xl, *ac, 0, 0, 0);
        *ac = __builtin_amdgcn_mfma_f32_32x32x16_f16(cl, xh, *ac, 0, 0, 0);
      }
    }
  };
  constexpr int NTRI = M * (M + 1) / 2;
  auto fetch_tri = [&](Stage& g, auto ec) {
    constexpr int e = decltype(ec)::value < NTRI ? decltype(ec)::value : NTRI - 1;
    constexpr int jj = tri_jj<M>(e), ii = tri_ii<M>(e);
    const bool act = ii < nb;
    fetch(g, act ? i0 + ii : i0, act ? i0 + jj : i0, i0 * 64);
  };

  Stage g;
  if (i0 > 0) {
    {  // X_0 of the first off-band column
      float v[16];
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = xsrc[(int64_t)u * kp];
      split8(v, xj_ptr(0), xk);
      split8(v + 8, xj_ptr(0), xk + 8);
    }
    // step (ii, j) stages rows 16 ii + 4 w .. + 3 of column j + 1
    fetch(g, i0, 0, 64 + 4 * w);
  } else {
    fetch_tri(g, std::integral_constant<int, 0>{});
  }
  // off-band columns j < i0: all M rows of the band each
  for (int j = 0; j < i0; j++) {
    static_for<0, M>([&](auto iic) {
      constexpr int ii = decltype(iic)::value;
      stash(g);
      stash_next(g, j + 1, 16 * ii + 4 * w);
      __syncthreads();
      // the next step: same column, the next column or the triangle's first
      if constexpr (ii + 1 < M) {
        fetch(g, i0 + ii + 1, j, (j + 1) * 64 + 16 * (ii + 1) + 4 * w);
      } else {
        const bool nxt = j + 1 < i0;
        fetch(g, i0, nxt ? j + 1 : i0, nxt ? min(j + 2, T - 1) * 64 + 4 * w : i0 * 64);
      }
      mm(accr[ii], xj_ptr(j), &accc[0], true);
      if constexpr (ii == M - 1) {  // column j done: its partial out
        float* pq = part + poff[d] + ((band_pbase(geo, b) + j) * (kp / 64) + blockIdx.z) * 4096;
#pragma unroll
        for (int r = 0; r < 16; r++) {
          pq[(w * 16 + r) * 64 + lane] = accc[0][r];
          accc[0][r] = 0.0f;
        }
      }
      __syncthreads();
    });
  }
  // the band's own lower triangle, (jj, ii) column-major; inactive steps (ii >= nb,
  // band 0 only) load a valid tile and do nothing else
  static_for<0, NTRI>([&](auto ec) {
    constexpr int e = decltype(ec)::value, jj = tri_jj<M>(e), ii = tri_ii<M>(e);
    const bool act = ii < nb;
    if (act) {
      stash(g);
      if constexpr (ii == jj) {  // the column's X_j is the diagonal tile's X_i
        split8(g.xi, xj_ptr(i0 + jj), xk);
        split8(g.xi + 8, xj_ptr(i0 + jj), xk + 8);
      }
    }
    __syncthreads();
    fetch_tri(g, std::integral_constant<int, e + 1>{});
    if (act) mm(accr[ii], xj_ptr(i0 + jj), &accc[jj], ii > jj);
    __syncthreads();
  });
  const int col = cb + (lane & 31);
  // the two scales applied one after the other: their product alone may underflow
  const float fa = alpha * sA[d], fx = sx[(int64_t)d * kp + c0 + col];
  float* Yd = Yf + xoff[d] + c0;
  static_for<0, M>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m < nb)
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int row = (i0 + m) * 64 + rb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        Yd[(int64_t)row * kp + col] = ((accr[m][r] + accc[m][r]) * fa) * fx;
      }
  });
}

// Y_j += alpha sA sx (sum of the partials of the bands below row block j, in
// ascending band order); one workgroup per (row block j, subdomain, column block),
// each thread 4 groups of 4 consecutive columns, every partial's loads issued
// before any is added
__global__ __launch_bounds__(256) void k_tgemm_h16b_reduce(const int* __restrict__ mrows,
                                                           const float* __restrict__ sA,
                                                           const int64_t* __restrict__ xoff,
                                                           const float* __restrict__ sx, int kp, float alpha,
                                                           const int* __restrict__ gate, float* __restrict__ Yf,
                                                           const float* __restrict__ part,
                                                           const int64_t* __restrict__ poff) {
  if (gate && *gate == 0) return;
  const int j = blockIdx.x, d = blockIdx.y;
  const BandGeom geo = band_geom(mrows[d]);
  if (j >= geo.T) return;
  const int bj = j < geo.nb0 ? 0 : 1 + (j - geo.nb0) / BM;
  if (bj >= geo.B - 1) return;
  const int c0 = blockIdx.z * 64;
  const float fa = alpha * sA[d];
  const float* pd = part + poff[d] + (int64_t)blockIdx.z * 4096;
  float* Yd = Yf + xoff[d] + c0;
  f4 s[4];
#pragma unroll
  for (int m = 0; m < 4; m++) s[m] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int b0 = bj + 1; b0 < geo.B; b0 += 4) {
    f4 v[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int bq = min(b0 + q, geo.B - 1);
      const f4* src = (const f4*)(pd + (band_pbase(geo, bq) + j) * (kp / 64) * 4096);
#pragma unroll
      for (int m = 0; m < 4; m++) v[q][m] = src[threadIdx.x + 256 * m];
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (b0 + q < geo.B)
#pragma unroll
        for (int m = 0; m < 4; m++) s[m] += v[q][m];
  }
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int e = 4 * (threadIdx.x + 256 * m);  // 4 consecutive lanes: 4 consecutive columns of one row
    const int wq = e >> 10, r = (e >> 6) & 15, lane = e & 63;
    const int row = j * 64 + (wq >> 1) * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int col = (wq & 1) * 32 + (lane & 31);
    const f4 fx = *(const f4*)(sx + (int64_t)d * kp + c0 + col);
    f4* y = (f4*)(Yd + (int64_t)row * kp + col);
    f4 o = *y;
#pragma unroll
    for (int u = 0; u < 4; u++) o[u] += (s[m][u] * fa) * fx[u];
    *y = o;
  }
}

int64_t tgemm_h16b_part_floats(int ns, int kp) {
  const BandGeom g = band_geom(ns);
  return band_pbase(g, g.B) * (kp / 64) * 4096;
}

hipError_t launch_tgemm_h16b(hipStream_t st, int nsub, int maxT, const int* ns, const float* A, const int64_t* aoff,
                             const float* sA, const float* X, const int64_t* xoff, float* sx, int kp, float alpha,
                             const int* gate, float* Yf, const int* xcd_order, int norder, int scales, float* part,
                             const int64_t* poff) {
  if (nsub <= 0 || maxT <= 0) return hipSuccess;
  if (kp <= 0 || kp % 64 || !part || !poff) return hipErrorInvalidValue;
  if (xcd_order && (norder < 8 || norder % 8)) return hipErrorInvalidValue;
  if (scales) {
    const hipError_t e = launch_kb_colscale(st, nsub, ns, xoff, kp, X, sx, gate);
    if (e != hipSuccess) return e;
  }
  const int groups = xcd_order ? norder / 8 : (nsub + 7) / 8;
  const int maxB = (maxT + BM - 1) / BM;
  hipLaunchKernelGGL((k_tgemm_h16b<BM>), dim3(8 * maxB * groups, 1, kp / 64), dim3(256), 0, st, nsub, maxB, xcd_order,
                     norder, ns, A, aoff, sA, X, xoff, (const float*)sx, kp, alpha, gate, Yf, part, poff);
  if (maxB > 1)
    hipLaunchKernelGGL(k_tgemm_h16b_reduce, dim3(maxT, nsub, kp / 64), dim3(256), 0, st, ns, sA, xoff,
                       (const float*)sx, kp, alpha, gate, Yf, (const float*)part, poff);
  return hipGetLastError();
}

}  // namespace hfls
