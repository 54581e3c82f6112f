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
// 16-byte loads and the transposing loads are all free of bank conflicts.  The
// band's M X blocks stay in LDS as hi/lo FP16 for the whole workgroup (80 KB of LDS
// with the tile: two workgroups per CU); X_j of an off-band column is held in
// registers as MFMA operands.  The tiles of the next M steps are in flight in
// registers (one per band row of an off-band column), and the next column's X_j
// is loaded M steps ahead.
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

// rows of LDS column c (64 halves each) are XORed in groups of 8 by
// f(c) = (16 (c mod 4) + 8 ((c / 4) mod 2)) ^ 16 ((c / 8) mod 2): the 4 columns one
// transposing load spans fall in 4 different 16-row blocks, and the 8 (or 16)
// consecutive columns that 8 (or 16) lanes of a 16-byte load address at once fall
// in 8 (or, with the column parity, 16) different bank groups
static __device__ __forceinline__ int swz8(int c) {
  return ((((c & 3) << 4) | (((c >> 2) & 1) << 3)) ^ (((c >> 3) & 1) << 4));
}
// position of tile element (row r, column c) in an LDS tile plane (column major)
static __device__ __forceinline__ int tpos(int c, int r) { return c * 64 + (r ^ swz8(c)); }

// position of X element (column n, row k) in an LDS X plane, swizzled like a tile
static __device__ __forceinline__ int xpos(int n, int k) { return n * 64 + (k ^ swz8(n)); }

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
  static_assert(M == 4, "two workgroups per CU: the tile and 4 X blocks in 80 KB of LDS");
  constexpr int XP = 4096;  // one 64 x 64 FP16 plane (halves)
  __shared__ __attribute__((aligned(16))) _Float16 lds[2 * XP * (M + 1)];
  _Float16* const Sh = lds;           // tile (tpos), hi
  _Float16* const Sl = lds + XP;      //             lo
  _Float16* const XB = lds + 2 * XP;  // the band's X block m: XB + 2 XP m (hi), + XP (lo), at xpos()

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
  // row block i0 + m: T X_j sums; column i0 + m (accc[0]: also the off-band partial).
  // accc[1..M-1] are first used in the in-band triangle and set to zero only there,
  // so their registers are free during the off-band columns
  floatx16 accr[M], accc[M];
  static_for<0, M>([&](auto mc) {
#pragma unroll
    for (int r = 0; r < 16; r++) accr[mc.value][r] = 0.0f;
  });
#pragma unroll
  for (int r = 0; r < 16; r++) accc[0][r] = 0.0f;

  const _Float16* Ad = (const _Float16*)(A + aoff[d]);
  const float* Xd = X + xoff[d] + c0;
  // tile staging: 16-byte units t and t + 256 of each plane (column (t >> 3) (+32),
  // rows (t & 7) 8 .. +7)
  const int tc = t >> 3, tr8 = (t & 7) * 8;
  const int sp0 = tpos(tc, tr8), sp1 = tpos(tc + 32, tr8);
  const _Float16* const tsrc = Ad + tc * 64 + tr8;
  // X operands: column xcol = cb + lane % 32, rows kk + ck .. +7
  const int xcol = cb + (lane & 31), ck = 8 * (lane >> 5);
  const float xinv_op = 1.0f / sx[(int64_t)d * kp + c0 + xcol];

  struct Stage {
    half8 h0, h1, l0, l1;
  };
  auto fetch = [&](Stage& g, int i, int j) {
    const int64_t ti = (int64_t)j * T - (int64_t)j * (j - 1) / 2 + (i - j);
    const _Float16* tile = tsrc + ti * 8192;
    g.h0 = *(const half8*)tile;
    g.h1 = *(const half8*)(tile + 2048);
    g.l0 = *(const half8*)(tile + 4096);
    g.l1 = *(const half8*)(tile + 6144);
  };
  auto stash = [&](const Stage& g) {
    *(half8*)(Sh + sp0) = g.h0;
    *(half8*)(Sh + sp1) = g.h1;
    *(half8*)(Sl + sp0) = g.l0;
    *(half8*)(Sl + sp1) = g.l1;
  };
  // X_j of an off-band column, as this lane's B operands (4 K steps x hi, lo): loaded
  // raw two steps ahead, split when the column starts
  struct XjRaw {
    float v[4][8];
  };
  struct XjOp {
    half8 h[4], l[4];
  };
  auto xj_load = [&](XjRaw& r, int col) {
    const float* xr = Xd + (int64_t)(col * 64 + ck) * kp + xcol;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int u = 0; u < 8; u++) r.v[q][u] = xr[(int64_t)(16 * q + u) * kp];
  };
  auto xj_split = [&](const XjRaw& r, XjOp& o) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const float x = r.v[q][u] * xinv_op;
        o.h[q][u] = (_Float16)x;
        o.l[q][u] = (_Float16)(x - (float)o.h[q][u]);
      }
  };
  // T X_j: lane group g4 = lane / 16 reads rows m0 + li of columns kg .. kg+7 with two
  // transposing loads (kg = kk + 8 (g4 / 2), m0 = rb + 16 (g4 % 2)); lane li addresses
  // column kg + li / 4 (+4), rows m0 + 4 (li % 4) .. +3
  const int g4 = lane >> 4, li = lane & 15;
  const int trow = rb + 16 * (g4 & 1) + 4 * (li & 3), tcol = 8 * (g4 >> 1) + (li >> 2);
  const int tro1 = tpos(tcol, trow), tro2 = tpos(tcol + 4, trow);  // + 64 kk
  const int ccol = rb + (lane & 31);  // T^T X_i: tile column rb + lane % 32, rows kk + ck .. +7
  auto tr4 = [](const _Float16* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4f16((__attribute__((address_space(3))) fp16x4*)p);
  };
  auto row_a = [&](int kk, half8& ah, half8& al) {
    ah = join8(tr4(Sh + tro1 + 64 * kk), tr4(Sh + tro2 + 64 * kk));
    al = join8(tr4(Sl + tro1 + 64 * kk), tr4(Sl + tro2 + 64 * kk));
  };
  auto mfma3 = [](floatx16& a, half8 ah, half8 al, half8 bh, half8 bl) {
    a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, a, 0, 0, 0);
    a = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, a, 0, 0, 0);
    a = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, a, 0, 0, 0);
  };
  auto col_prod = [&](floatx16& a, const _Float16* xi, int kk) {  // a += (T^T X_i), K step kk
    const int o = tpos(ccol, kk + ck), xo = xpos(xcol, kk + ck);
    mfma3(a, *(const half8*)(Sh + o), *(const half8*)(Sl + o), *(const half8*)(xi + xo),
          *(const half8*)(xi + XP + xo));
  };
  constexpr int NTRI = M * (M + 1) / 2;
  // step e of the in-band triangle -> tile; inactive (row >= nb, band 0 only) steps
  // load a valid tile and do nothing else
  auto fetch_tri = [&](Stage& g, auto ec) {
    constexpr int e = decltype(ec)::value < NTRI ? decltype(ec)::value : NTRI - 1;
    constexpr int jj = tri_jj<M>(e), ii = tri_ii<M>(e);
    const bool act = ii < nb;
    fetch(g, act ? i0 + ii : i0, act ? i0 + jj : i0);
  };

  // the band's X blocks: thread column t % 64, rows 16 (t / 64) .. +15 of each
  {
    const int xn = t & 63, xk = 16 * w;
    const float xinv = 1.0f / sx[(int64_t)d * kp + c0 + xn];
    for (int m = 0; m < nb; m++) {
      float v[16];
      const float* xr = Xd + (int64_t)((i0 + m) * 64 + xk) * kp + xn;
#pragma unroll
      for (int u = 0; u < 16; u++) v[u] = xr[(int64_t)u * kp];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        half8 hh, ll;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const float x = v[8 * h + u] * xinv;
          hh[u] = (_Float16)x;
          ll[u] = (_Float16)(x - (float)hh[u]);
        }
        _Float16* blk = XB + 2 * XP * m;
        *(half8*)(blk + xpos(xn, xk + 8 * h)) = hh;
        *(half8*)(blk + XP + xpos(xn, xk + 8 * h)) = ll;
      }
    }
  }
  // tiles of steps s .. s + M - 1 in flight, slot s mod M: with M steps per off-band
  // column the slot of a step is its row ii in the band, and the in-band triangle
  // starts at step M i0, so every slot index is a compile-time constant
  Stage g[M];
  XjRaw xraw;
  XjOp xop;
  if (i0 > 0) {
    static_for<0, M>([&](auto uc) { fetch(g[uc.value], i0 + uc.value, 0); });
    xj_load(xraw, 0);
  } else {
    static_for<0, M>([&](auto uc) { fetch_tri(g[uc.value], uc); });
  }
  // off-band columns j < i0: the band's M rows each
  for (int j = 0; j < i0; j++) {
    static_for<0, M>([&](auto iic) {
      constexpr int ii = decltype(iic)::value;  // step M j + ii, slot ii
      if constexpr (ii == 0) xj_split(xraw, xop);  // this column's X_j, loaded M steps ago
      stash(g[ii]);
      __syncthreads();
      {  // the tile M steps ahead: row ii of the next column, or step ii of the triangle
        constexpr int tj = tri_jj<M>(ii), tii = tri_ii<M>(ii);
        const bool nxt = j + 1 < i0;
        fetch(g[ii], nxt ? i0 + ii : i0 + tii, nxt ? j + 1 : i0 + tj);
      }
      if constexpr (ii == M - 2) xj_load(xraw, min(j + 1, T - 1));  // the next column's X_j, two steps ahead
      __builtin_amdgcn_iglp_opt(0);  // interleave the LDS operand reads with the MFMAs
      const _Float16* xi = XB + 2 * XP * ii;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        half8 ah, al;
        row_a(16 * q, ah, al);
        mfma3(accr[ii], ah, al, xop.h[q], xop.l[q]);
        col_prod(accc[0], xi, 16 * q);
      }
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
  static_for<1, M>([&](auto mc) {
#pragma unroll
    for (int r = 0; r < 16; r++) accc[mc.value][r] = 0.0f;
  });
  // the band's own lower triangle, (jj, ii) column-major
  static_for<0, NTRI>([&](auto ec) {
    constexpr int e = decltype(ec)::value, jj = tri_jj<M>(e), ii = tri_ii<M>(e);
    const bool act = ii < nb;
    constexpr int q = e % M;  // step M i0 + e
    if (act) stash(g[q]);
    __syncthreads();
    if constexpr (e + M < NTRI) fetch_tri(g[q], std::integral_constant<int, e + M>{});
    if (act) {
      const _Float16* xj = XB + 2 * XP * jj;
      __builtin_amdgcn_iglp_opt(0);  // interleave the LDS operand reads with the MFMAs
      const _Float16* xi = XB + 2 * XP * ii;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        half8 ah, al;
        row_a(16 * q, ah, al);
        const int xo = xpos(xcol, 16 * q + ck);
        mfma3(accr[ii], ah, al, *(const half8*)(xj + xo), *(const half8*)(xj + XP + xo));
        if constexpr (ii > jj) col_prod(accc[jj], xi, 16 * q);
      }
    }
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
