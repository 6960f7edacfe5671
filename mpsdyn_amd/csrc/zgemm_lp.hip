// Precision forms of the matvec on gfx950: operands packed once into low precision, and the
// complex BRGEMM that reads them packed.
//
// Packed layout.  A matrix Op[row][k] of one batch (k the contraction index) is stored row
// after row; a row holds Kp = K rounded up to 32 complex values (zeros past K) as 64-byte
// groups of 8 consecutive k:
//   FP32    { re[k..k+3], im[k..k+3], re[k+4..k+7], im[k+4..k+7] }             (float)
//   BF16X3  { re_hi[k..k+7], re_lo[k..k+7], im_hi[k..k+7], im_lo[k..k+7] }     (bf16)
// with x = hi + lo, hi = bf16(x), lo = bf16(x - hi), both rounded to nearest even from FP32.
// Both forms take 8 bytes per complex value, half of FP64, and a 32-deep K stage of a row is
// 256 contiguous bytes.
//
// The environments L and R of a matvec are packed once per handle; x is packed per call, the
// first stage writes T1 in FP32, the MPO mix packs T2 for the last stage, and the last stage's
// per-MPO-bond partial products (FP32) are summed into the FP64 result with the offset term.
//
// BRGEMM: a workgroup of 2 x 2 waves owns a (32 BM) x (32 BN) output tile, each wave BM x BN
// blocks of 16 x 16.  K stages of 32 are double-buffered in LDS with the rows as they are in
// HBM (pitch 272 bytes, so the 16 rows a lane group reads start 4 banks apart); lane group
// q = lane / 16 reads the 64-byte group k = 8q .. 8q+7 of its row as four 16-byte reads:
//  - BF16X3: that group is the fragment of v_mfma_f32_16x16x32_bf16 (A[m][k] from lane
//    m + 16 (k / 8), 8 consecutive k per lane), one plane per 16 bytes.  A complex block is the
//    four real products, each as hi*hi + hi*lo + lo*hi (lo*lo, ~2^-16 relative, dropped):
//    12 MFMAs per 16 x 16 x 32 block.
//  - FP32: v_mfma_f32_16x16x4f32 takes k from lane group q at each of its 4-deep steps; step j
//    of the stage's 8 gives lane group q the value k = 8q + j, the same permutation for A and
//    B, so every k of the stage is summed once.  The complex product is the three-product form
//    (ar br, ai bi, (ar + ai)(br + bi)): 24 MFMAs per 16 x 16 x 32 block.
// Result element (m, n) of both MFMAs is in lane n + 16 (m / 4), register m % 4.

#include "mpsd_internal.h"

namespace mpsd {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int kT = 256;
constexpr int kPitch = 17;  // uint4 per LDS row: 16 for the 32-deep stage + 1 of padding
constexpr int kGroupM = 8;  // row-tiles per tile group (split-BF16 and FP32 launches)

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// Pack 8 consecutive complex values into one 64-byte group (out[0..3], 16 bytes each).
template <int PREC>
__device__ __forceinline__ void pack8(const float re[8], const float im[8], uint4 out[4]) {
  if constexpr (PREC == MPSD_PREC_FP32) {
    out[0] = make_uint4(__float_as_uint(re[0]), __float_as_uint(re[1]), __float_as_uint(re[2]), __float_as_uint(re[3]));
    out[1] = make_uint4(__float_as_uint(im[0]), __float_as_uint(im[1]), __float_as_uint(im[2]), __float_as_uint(im[3]));
    out[2] = make_uint4(__float_as_uint(re[4]), __float_as_uint(re[5]), __float_as_uint(re[6]), __float_as_uint(re[7]));
    out[3] = make_uint4(__float_as_uint(im[4]), __float_as_uint(im[5]), __float_as_uint(im[6]), __float_as_uint(im[7]));
  } else {
    unsigned rh[4], rl[4], ih[4], il[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const unsigned short h0 = f2bf(re[2 * j]), h1 = f2bf(re[2 * j + 1]);
      const unsigned short l0 = f2bf(re[2 * j] - bf2f(h0)), l1 = f2bf(re[2 * j + 1] - bf2f(h1));
      const unsigned short g0 = f2bf(im[2 * j]), g1 = f2bf(im[2 * j + 1]);
      const unsigned short m0 = f2bf(im[2 * j] - bf2f(g0)), m1 = f2bf(im[2 * j + 1] - bf2f(g1));
      rh[j] = (unsigned)h0 | ((unsigned)h1 << 16);
      rl[j] = (unsigned)l0 | ((unsigned)l1 << 16);
      ih[j] = (unsigned)g0 | ((unsigned)g1 << 16);
      il[j] = (unsigned)m0 | ((unsigned)m1 << 16);
    }
    out[0] = make_uint4(rh[0], rh[1], rh[2], rh[3]);
    out[1] = make_uint4(rl[0], rl[1], rl[2], rl[3]);
    out[2] = make_uint4(ih[0], ih[1], ih[2], ih[3]);
    out[3] = make_uint4(il[0], il[1], il[2], il[3]);
  }
}

// Four values of pack8's layout: a_r = 8 g + 4 h + j (j < 4) of one row, stored at dst (the
// row's four 16-byte units of group g), the half h of each unit that pack8 would write.
template <int PREC>
__device__ __forceinline__ void store_pack4(const float re[4], const float im[4], int h, uint4* dst) {
  if constexpr (PREC == MPSD_PREC_FP32) {
    dst[2 * h] = make_uint4(__float_as_uint(re[0]), __float_as_uint(re[1]), __float_as_uint(re[2]),
                            __float_as_uint(re[3]));
    dst[2 * h + 1] = make_uint4(__float_as_uint(im[0]), __float_as_uint(im[1]), __float_as_uint(im[2]),
                                __float_as_uint(im[3]));
  } else {
    unsigned rh[2], rl[2], ih[2], il[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const unsigned short h0 = f2bf(re[2 * j]), h1 = f2bf(re[2 * j + 1]);
      const unsigned short l0 = f2bf(re[2 * j] - bf2f(h0)), l1 = f2bf(re[2 * j + 1] - bf2f(h1));
      const unsigned short g0 = f2bf(im[2 * j]), g1 = f2bf(im[2 * j + 1]);
      const unsigned short m0 = f2bf(im[2 * j] - bf2f(g0)), m1 = f2bf(im[2 * j + 1] - bf2f(g1));
      rh[j] = (unsigned)h0 | ((unsigned)h1 << 16);
      rl[j] = (unsigned)l0 | ((unsigned)l1 << 16);
      ih[j] = (unsigned)g0 | ((unsigned)g1 << 16);
      il[j] = (unsigned)m0 | ((unsigned)m1 << 16);
    }
    reinterpret_cast<uint2*>(dst + 0)[h] = make_uint2(rh[0], rh[1]);
    reinterpret_cast<uint2*>(dst + 1)[h] = make_uint2(rl[0], rl[1]);
    reinterpret_cast<uint2*>(dst + 2)[h] = make_uint2(ih[0], ih[1]);
    reinterpret_cast<uint2*>(dst + 3)[h] = make_uint2(il[0], il[1]);
  }
}

// FP64 strided source -> packed operand.  One thread per (batch, row, group of 8 k, half h of
// the group), h fastest, then rows (the environments are row-contiguous in HBM).  Half groups
// rather than whole ones: twice the threads, which made the per-call pack of x 0.7-2.6 us
// faster at D = 64-512 (the same conversions, the same packed bits).
template <int PREC>
__global__ void __launch_bounds__(kT) lp_pack_kernel(const LpPackDesc p) {
  const int64_t ng = p.Kp / 8;
  const int64_t total = (int64_t)p.nb * p.rows * ng * 2;
  const bool row_fast = p.srow == 1;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int64_t t2 = t >> 1;
    const int h = (int)(t & 1);
    const int b = (int)(t2 / ((int64_t)p.rows * ng));
    const int64_t r = t2 - (int64_t)b * p.rows * ng;
    const int row = row_fast ? (int)(r % p.rows) : (int)(r / ng);
    const int g = row_fast ? (int)(r / p.rows) : (int)(r % ng);
    const double2* src = p.src + (int64_t)b * p.sb + (int64_t)row * p.srow;
    float re[4], im[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 8 * g + 4 * h + j;
      const double2 v = k < p.K ? src[(int64_t)k * p.sk] : double2{0.0, 0.0};
      re[j] = (float)v.x;
      im[j] = (float)v.y;
    }
    uint4* dst = reinterpret_cast<uint4*>(p.dst) + (((int64_t)b * p.rows + row) * p.Kp + 8 * g) / 2;
    store_pack4<PREC>(re, im, h, dst);
  }
}

// T2 = MPO mix of T1 (FP32), packed as the last stage's A operand: batch o = s' + d be, rows b,
// k = a_r.  T1[b + Db i + Db nin a_r] (i = al + wl s), mix[o + nout i] (FP64, rounded to FP32).
// One thread per (b, group of 8 a_r, pair of outputs o), b fastest.
template <int PREC>
__global__ void __launch_bounds__(kT) lp_mix_kernel(const float2* __restrict__ T1, void* __restrict__ T2p,
                                                    const double2* __restrict__ mix, int Db, int nin, int nout,
                                                    int Dkr, int Kp) {
  __shared__ float2 wm[16 * 16];
  for (int i = threadIdx.x; i < nin * nout; i += kT) {
    const double2 v = mix[i];
    wm[i] = float2{(float)v.x, (float)v.y};
  }
  __syncthreads();
  const int ng = Kp / 8, npair = (nout + 1) / 2;
  const int64_t total = (int64_t)Db * ng * npair;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int b = (int)(t % Db);
    const int64_t rest = t / Db;
    const int g = (int)(rest % ng);
    const int o0 = 2 * (int)(rest / ng);
    float re[2][8], im[2][8];
#pragma unroll
    for (int o = 0; o < 2; ++o)
#pragma unroll
      for (int j = 0; j < 8; ++j) re[o][j] = im[o][j] = 0.f;
    const bool two = o0 + 1 < nout;
    for (int i = 0; i < nin; ++i) {
      const float2 w0 = wm[o0 + nout * i];
      const float2 w1 = two ? wm[o0 + 1 + nout * i] : float2{0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int ar = 8 * g + j;
        const float2 v = ar < Dkr ? T1[b + (int64_t)Db * i + (int64_t)Db * nin * ar] : float2{0.f, 0.f};
        re[0][j] = fmaf(w0.x, v.x, fmaf(-w0.y, v.y, re[0][j]));
        im[0][j] = fmaf(w0.x, v.y, fmaf(w0.y, v.x, im[0][j]));
        re[1][j] = fmaf(w1.x, v.x, fmaf(-w1.y, v.y, re[1][j]));
        im[1][j] = fmaf(w1.x, v.y, fmaf(w1.y, v.x, im[1][j]));
      }
    }
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      if (o == 1 && !two) break;
      uint4 out[4];
      pack8<PREC>(re[o], im[o], out);
      // row-interleaved: group g of row b at (g Db + b) 4 (16-byte units) within output o
      uint4* dst = reinterpret_cast<uint4*>(T2p) + (int64_t)(o0 + o) * Db * Kp / 2 + ((int64_t)g * Db + b) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) dst[u] = out[u];
    }
  }
}

// lp_mix_kernel's sums with one output per grid row (o = blockIdx.y) and half a group of a_r per
// thread: 2 nout / 2 times its threads, for the small grids where each of its threads would
// carry 16 sums.  The same fmaf sequence per value, so the same bits.
template <int PREC>
__global__ void __launch_bounds__(kT) lp_mix_split_kernel(const float2* __restrict__ T1, void* __restrict__ T2p,
                                                          const double2* __restrict__ mix, int Db, int nin,
                                                          int nout, int Dkr, int Kp) {
  __shared__ float2 wm[256];
  const int o = blockIdx.y;
  for (int i = threadIdx.x; i < nin; i += kT) {
    const double2 v = mix[o + nout * i];
    wm[i] = float2{(float)v.x, (float)v.y};
  }
  __syncthreads();
  const int ng = Kp / 8;
  const int64_t total = (int64_t)Db * ng * 2;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int b = (int)(t % Db);
    const int64_t rest = t / Db;
    const int h = (int)(rest & 1);
    const int g = (int)(rest >> 1);
    float re[4], im[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) re[j] = im[j] = 0.f;
    for (int i = 0; i < nin; ++i) {
      const float2 w0 = wm[i];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int ar = 8 * g + 4 * h + j;
        const float2 v = ar < Dkr ? T1[b + (int64_t)Db * i + (int64_t)Db * nin * ar] : float2{0.f, 0.f};
        re[j] = fmaf(w0.x, v.x, fmaf(-w0.y, v.y, re[j]));
        im[j] = fmaf(w0.x, v.y, fmaf(w0.y, v.x, im[j]));
      }
    }
    uint4* dst = reinterpret_cast<uint4*>(T2p) + (int64_t)o * Db * Kp / 2 + ((int64_t)g * Db + b) * 4;
    store_pack4<PREC>(re, im, h, dst);
  }
}

// y[i] = sum_{q < np} P[q * stride + i] (FP32 partials, summed in FP64 in a fixed order) + beta x[i]
__global__ void __launch_bounds__(kT) lp_sum_kernel(const float2* __restrict__ P, int np, int64_t stride,
                                                    const double2* __restrict__ x, double beta,
                                                    double2* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  double re = 0.0, im = 0.0;
  for (int q = 0; q < np; ++q) {
    const float2 v = P[(int64_t)q * stride + i];
    re += (double)v.x;
    im += (double)v.y;
  }
  if (x) {
    const double2 v = x[i];
    re = fma(beta, v.x, re);
    im = fma(beta, v.y, im);
  }
  y[i] = double2{re, im};
}

// Workgroup -> (tile, batch).  Ids are regrouped so each XCD owns a contiguous range; within it
// the first batch index (z % zdiv: the physical index s or s' of the stages) is fastest, so the
// zdiv workgroups of one tile run side by side on one XCD: they share the common operand's
// panel in its L2 and write the interleaved halves of the same output lines.  Tiles are walked
// in groups of `group` row-tiles; the second batch index (z / zdiv) is slowest.
__device__ __forceinline__ void lp_tile_of(int gm, int gn, int zdiv, int group, int& tm_i, int& tn_i, int& z) {
  const int total = gridDim.x;
  int id = blockIdx.x;
  if ((total & 7) == 0) id = (id & 7) * (total >> 3) + (id >> 3);  // contiguous range per XCD
  const int z0 = id % zdiv;
  const int rest = id / zdiv;
  const int per_z = gm * gn;
  const int z1 = rest / per_z;
  const int r = rest - z1 * per_z;
  z = z0 + zdiv * z1;
  const int in_group = group * gn;
  const int g = r / in_group;
  const int first_m = g * group;
  const int gsize = min(gm - first_m, group);
  const int rr = r - g * in_group;
  tm_i = first_m + rr % gsize;
  tn_i = rr / gsize;
}

__device__ __forceinline__ f4 mma_bf(u4v a, u4v b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

template <int PREC, int BM, int BN>
__global__ void __launch_bounds__(kT) lp_brgemm_kernel(const LpDesc p, int gm, int gn) {
  constexpr int TM = 32 * BM, TN = 32 * BN;
  constexpr int AL = TM * 16 / kT;  // 16-byte loads per thread and stage
  constexpr int BL = TN * 16 / kT;
  constexpr bool BF = PREC == MPSD_PREC_BF16X3;
  constexpr int NACC = BF ? 2 : 3;
  // u4v rather than HIP's uint4 struct throughout, so the staging arrays stay in registers
  __shared__ u4v As[2][TM][kPitch];
  __shared__ u4v Bs[2][TN][kPitch];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tm_i, tn_i, z;
  lp_tile_of(gm, gn, p.zdiv, p.group > 0 ? p.group : kGroupM, tm_i, tn_i, z);
  const int m0 = tm_i * TM, n0 = tn_i * TN;
  const int z0 = z % p.zdiv, z1 = z / p.zdiv;
  const u4v* Ab = reinterpret_cast<const u4v*>(p.A) + (z0 * p.sA0 + z1 * p.sA1) / 2;
  const u4v* Bb = reinterpret_cast<const u4v*>(p.B) + (z0 * p.sB0 + z1 * p.sB1) / 2;
  const int64_t rowq = p.Kp / 2;  // uint4 per packed row
  const int c16 = tid & 15;       // 16-byte chunk of a row's stage
  const int r16 = tid >> 4;       // row within a group of 16
  // A row-interleaved (a_il): the 4-chunk groups G of all rows in turn, at (G M + row) 4 + chunk;
  // a wave loads one group of 16 rows, 1 KB contiguous
  const int ar16 = p.a_il ? (tid >> 2) & 15 : r16;
  const int ac16 = p.a_il ? 4 * (tid >> 6) + (tid & 3) : c16;
  const int64_t astep = p.a_il ? 16 * (int64_t)p.M : 16;

  f4 acc[NACC][BM][BN];
#pragma unroll
  for (int c = 0; c < NACC; ++c)
#pragma unroll
    for (int i = 0; i < BM; ++i)
#pragma unroll
      for (int j = 0; j < BN; ++j) acc[c][i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // rows past M or N load a valid row (the last one): they only feed outputs that are dropped
  const u4v* arow[AL];
  const u4v* brow[BL];
#pragma unroll
  for (int e = 0; e < AL; ++e) {
    const int64_t row = min(m0 + ar16 + 16 * e, p.M - 1);
    arow[e] = Ab + (p.a_il ? ((ac16 >> 2) * (int64_t)p.M + row) * 4 + (ac16 & 3) : row * rowq + ac16);
  }
#pragma unroll
  for (int e = 0; e < BL; ++e) brow[e] = Bb + (int64_t)min(n0 + r16 + 16 * e, p.N - 1) * rowq + c16;
  u4v ra[AL], rb[BL];
  auto load = [&](int t) {
#pragma unroll
    for (int e = 0; e < AL; ++e) ra[e] = arow[e][astep * t];
#pragma unroll
    for (int e = 0; e < BL; ++e) rb[e] = brow[e][16 * t];
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < AL; ++e) As[buf][ar16 + 16 * e][ac16] = ra[e];
#pragma unroll
    for (int e = 0; e < BL; ++e) Bs[buf][r16 + 16 * e][c16] = rb[e];
  };

  const int nst = p.Kp / 32;
  const int lr = lane & 15, q = lane >> 4;
  if (nst > 0) {
    load(0);
    store(0);
    __syncthreads();
  }
  for (int t = 0; t < nst; ++t) {
    const int cur = t & 1;
    if (t + 1 < nst) load(t + 1);
    u4v a[BM][4], b[BN][4];
#pragma unroll
    for (int i = 0; i < BM; ++i)
#pragma unroll
      for (int u = 0; u < 4; ++u) a[i][u] = As[cur][wm * (TM / 2) + i * 16 + lr][4 * q + u];
#pragma unroll
    for (int j = 0; j < BN; ++j)
#pragma unroll
      for (int u = 0; u < 4; ++u) b[j][u] = Bs[cur][wn * (TN / 2) + j * 16 + lr][4 * q + u];
    if constexpr (BF) {
#pragma unroll
      for (int i = 0; i < BM; ++i) {
        const u4v sgn = u4v{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
        const u4v rh = a[i][0], rl = a[i][1], ih = a[i][2], il = a[i][3];
        const u4v nih = ih ^ sgn, nil = il ^ sgn;
#pragma unroll
        for (int j = 0; j < BN; ++j) {
          const u4v brh = b[j][0], brl = b[j][1], bih = b[j][2], bil = b[j][3];
          f4 cr = acc[0][i][j], ci = acc[1][i][j];
          cr = mma_bf(rh, brh, cr);
          ci = mma_bf(rh, bih, ci);
          cr = mma_bf(rh, brl, cr);
          ci = mma_bf(rh, bil, ci);
          cr = mma_bf(rl, brh, cr);
          ci = mma_bf(rl, bih, ci);
          cr = mma_bf(nih, bih, cr);
          ci = mma_bf(ih, brh, ci);
          cr = mma_bf(nih, bil, cr);
          ci = mma_bf(ih, brl, ci);
          cr = mma_bf(nil, bih, cr);
          ci = mma_bf(il, brh, ci);
          acc[0][i][j] = cr;
          acc[1][i][j] = ci;
        }
      }
    } else {
      // step s of 8: k = 8q + s; element s of the group is (s < 4 ? u0 / u1 : u2 / u3)[s % 4]
      float ar[BM][8], ai[BM][8], as[BM][8], br[BN][8], bi[BN][8], bs[BN][8];
#pragma unroll
      for (int i = 0; i < BM; ++i)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const u4v R = a[i][s < 4 ? 0 : 2], I = a[i][s < 4 ? 1 : 3];
          ar[i][s] = __uint_as_float(R[s & 3]);
          ai[i][s] = __uint_as_float(I[s & 3]);
          as[i][s] = ar[i][s] + ai[i][s];
        }
#pragma unroll
      for (int j = 0; j < BN; ++j)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const u4v R = b[j][s < 4 ? 0 : 2], I = b[j][s < 4 ? 1 : 3];
          br[j][s] = __uint_as_float(R[s & 3]);
          bi[j][s] = __uint_as_float(I[s & 3]);
          bs[j][s] = br[j][s] + bi[j][s];
        }
#pragma unroll
      for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int i = 0; i < BM; ++i)
#pragma unroll
          for (int j = 0; j < BN; ++j) {
            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i][s], br[j][s], acc[0][i][j], 0, 0, 0);
            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i][s], bi[j][s], acc[1][i][j], 0, 0, 0);
            acc[2][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(as[i][s], bs[j][s], acc[2][i][j], 0, 0, 0);
          }
    }
    if (t + 1 < nst) store(cur ^ 1);
    __syncthreads();
  }

  float2* Cb = p.C + z0 * p.sC0 + z1 * p.sC1;
#pragma unroll
  for (int i = 0; i < BM; ++i)
#pragma unroll
    for (int j = 0; j < BN; ++j) {
      const int n = n0 + wn * (TN / 2) + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * (TM / 2) + i * 16 + 4 * q + r;
        if (m < p.M && n < p.N) {
          float2 v;
          if constexpr (BF) {
            v = float2{acc[0][i][j][r], acc[1][i][j][r]};
          } else {
            const float p1 = acc[0][i][j][r], p2 = acc[1][i][j][r], p3 = acc[2][i][j][r];
            v = float2{p1 - p2, p3 - (p1 + p2)};
          }
          Cb[m * p.sCm + n * p.sCn] = v;
        }
      }
    }
}

// ---------------------------------------------------------------- FP64 on packed operands
//
// The FP64 matvec on the same packed rows, 16 bytes per complex value and 16 k per stage:
// within each stage the value k = 16 t + 4 u + q sits in slot 4 q + u, so lane group q reads
// the four values it feeds to MFMA steps u = 0 .. 3 as four contiguous 16-byte reads.  Step u
// of every 16 x 16 block is v_mfma_f64_16x16x4f64 over k = 16 t + 4 u .. 16 t + 4 u + 3 with
// lane group q on k = 16 t + 4 u + q, in the three-product form, in ascending k: exactly the
// MFMA sequence of zbrgemm_f64 and of the one-workgroup chain, so the results are the same bits.

// FP64 strided source -> packed FP64 (permuted within each 16-deep stage).  One thread per
// (batch, row, stage, lane group); rows fastest when they are contiguous in the source (the
// environments), the (stage, lane group) fastest otherwise (x, whose k is the fast index).
__global__ void __launch_bounds__(kT) pk64_pack_kernel(const LpPackDesc p) {
  const int64_t nq = p.Kp / 4;  // (stage, lane group) pairs per row
  const int64_t total = (int64_t)p.nb * p.rows * nq;
  const bool row_fast = p.srow == 1;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int b = (int)(t / ((int64_t)p.rows * nq));
    const int64_t r = t - (int64_t)b * p.rows * nq;
    const int row = row_fast ? (int)(r % p.rows) : (int)(r / nq);
    const int sq = row_fast ? (int)(r / p.rows) : (int)(r % nq);
    const int st = sq >> 2, q = sq & 3;
    const double2* src = p.src + (int64_t)b * p.sb + (int64_t)row * p.srow;
    double2* dst = reinterpret_cast<double2*>(p.dst) + ((int64_t)b * p.rows + row) * p.Kp + 16 * st + 4 * q;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = 16 * st + 4 * u + q;
      dst[u] = k < p.K ? src[(int64_t)k * p.sk] : double2{0.0, 0.0};
    }
  }
}

// T2 = MPO mix of T1 (FP64, zmix_f64's arithmetic element by element), packed FP64 row-interleaved
// as the last stage's A operand.  One thread per T2 element of every output o (as zmix_f64),
// ordered (slot u, row b, stage group) so that a wave reads 16 consecutive b of four a_r and
// writes 1 KB contiguous per output.
constexpr int kMixMax = 16;
constexpr unsigned kMixSplitBelow = 512;  // one-row grids smaller than this split by output
constexpr unsigned kLpMixSplitBelow = 128;  // the same for the low-precision mix
__global__ void __launch_bounds__(kT) pk64_mix_kernel(const double2* __restrict__ T1, double2* __restrict__ T2p,
                                                      const double2* __restrict__ mix, int Db, int nin, int nout,
                                                      int Dkr, int Kp) {
  const int64_t total = (int64_t)Db * Kp;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int u = (int)(t & 3);
    const int64_t rest = t >> 2;
    const int b = (int)(rest % Db);
    const int G = (int)(rest / Db);  // group 4 st + q of the packed row
    const int ar = 16 * (G >> 2) + 4 * u + (G & 3);
    double2 acc[kMixMax];
#pragma unroll
    for (int o = 0; o < kMixMax; ++o) acc[o] = double2{0.0, 0.0};
    if (ar < Dkr) {
      for (int i = 0; i < nin; ++i) {
        const double2 v = T1[b + (int64_t)Db * i + (int64_t)Db * nin * ar];
#pragma unroll
        for (int o = 0; o < kMixMax; ++o) {
          if (o < nout) {
            const double2 w = mix[o + nout * i];
            acc[o].x = fma(w.x, v.x, fma(-w.y, v.y, acc[o].x));
            acc[o].y = fma(w.x, v.y, fma(w.y, v.x, acc[o].y));
          }
        }
      }
    }
    double2* dst = T2p + ((int64_t)G * Db + b) * 4 + u;
#pragma unroll
    for (int o = 0; o < kMixMax; ++o)
      if (o < nout) dst[(int64_t)o * Db * Kp] = acc[o];
  }
}

// The same sums with one output per grid row (o = blockIdx.y): nout times the workgroups and a
// sixth of the work per thread, for shapes whose one-row grid leaves most CUs idle (the T1
// values are read once per output, from L2).  Same operations per output, same bits.
__global__ void __launch_bounds__(kT) pk64_mix_split_kernel(const double2* __restrict__ T1,
                                                            double2* __restrict__ T2p,
                                                            const double2* __restrict__ mix, int Db, int nin,
                                                            int nout, int Dkr, int Kp) {
  const int o = blockIdx.y;
  const int64_t total = (int64_t)Db * Kp;
  for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < total; t += (int64_t)gridDim.x * kT) {
    const int u = (int)(t & 3);
    const int64_t rest = t >> 2;
    const int b = (int)(rest % Db);
    const int G = (int)(rest / Db);
    const int ar = 16 * (G >> 2) + 4 * u + (G & 3);
    double2 acc{0.0, 0.0};
    if (ar < Dkr) {
      for (int i = 0; i < nin; ++i) {
        const double2 v = T1[b + (int64_t)Db * i + (int64_t)Db * nin * ar];
        const double2 w = mix[o + nout * i];
        acc.x = fma(w.x, v.x, fma(-w.y, v.y, acc.x));
        acc.y = fma(w.x, v.y, fma(w.y, v.x, acc.y));
      }
    }
    T2p[(int64_t)o * Db * Kp + ((int64_t)G * Db + b) * 4 + u] = acc;
  }
}

template <int BM, int BN>
__global__ void __launch_bounds__(kT) pk64_brgemm_kernel(const LpDesc p, int gm, int gn) {
  constexpr int TM = 32 * BM, TN = 32 * BN;
  constexpr int AL = TM * 16 / kT;
  constexpr int BL = TN * 16 / kT;
  __shared__ u4v As[2][TM][kPitch];
  __shared__ u4v Bs[2][TN][kPitch];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tm_i, tn_i, z;
  lp_tile_of(gm, gn, p.zdiv, p.group > 0 ? p.group : kGroupM, tm_i, tn_i, z);
  const int m0 = tm_i * TM, n0 = tn_i * TN;
  const int z0 = z % p.zdiv, z1 = z / p.zdiv;
  const u4v* Ab = reinterpret_cast<const u4v*>(p.A) + (z0 * p.sA0 + z1 * p.sA1);
  const u4v* Bb = reinterpret_cast<const u4v*>(p.B) + (z0 * p.sB0 + z1 * p.sB1);
  const int64_t rowq = p.Kp;  // 16-byte values per packed row
  const int c16 = tid & 15, r16 = tid >> 4;
  const int ar16 = p.a_il ? (tid >> 2) & 15 : r16;
  const int ac16 = p.a_il ? 4 * (tid >> 6) + (tid & 3) : c16;
  const int64_t astep = p.a_il ? 16 * (int64_t)p.M : 16;

  d4 acc[3][BM][BN];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < BM; ++i)
#pragma unroll
      for (int j = 0; j < BN; ++j) acc[c][i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const u4v* arow[AL];
  const u4v* brow[BL];
#pragma unroll
  for (int e = 0; e < AL; ++e) {
    const int64_t row = min(m0 + ar16 + 16 * e, p.M - 1);
    arow[e] = Ab + (p.a_il ? ((ac16 >> 2) * (int64_t)p.M + row) * 4 + (ac16 & 3) : row * rowq + ac16);
  }
  // B read in place (b_ip: rows at sBn, k at sBk, the value of slot c = 4 q + u being
  // k = 16 t + 4 u + q, zero past Kb) or packed
  const int bk = 4 * (c16 & 3) + (c16 >> 2);
  const int64_t bstep = p.b_ip ? 16 * p.sBk : 16;
#pragma unroll
  for (int e = 0; e < BL; ++e) {
    const int64_t row = min(n0 + r16 + 16 * e, p.N - 1);
    brow[e] = Bb + (p.b_ip ? row * p.sBn + bk * p.sBk : row * rowq + c16);
  }
  u4v ra[AL], rb[BL];
  auto load = [&](int t) {
#pragma unroll
    for (int e = 0; e < AL; ++e) ra[e] = arow[e][astep * t];
    const bool bin = !p.b_ip || 16 * t + bk < p.Kb;
#pragma unroll
    for (int e = 0; e < BL; ++e) rb[e] = bin ? brow[e][bstep * t] : u4v{0u, 0u, 0u, 0u};
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < AL; ++e) As[buf][ar16 + 16 * e][ac16] = ra[e];
#pragma unroll
    for (int e = 0; e < BL; ++e) Bs[buf][r16 + 16 * e][c16] = rb[e];
  };

  const int nst = p.Kp / 16;
  const int lr = lane & 15, q = lane >> 4;
  if (nst > 0) {
    load(0);
    store(0);
    __syncthreads();
  }
  for (int t = 0; t < nst; ++t) {
    const int cur = t & 1;
    if (t + 1 < nst) load(t + 1);
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      double ar[BM], ai[BM], a3[BM], br[BN], bi[BN], b3[BN];
#pragma unroll
      for (int i = 0; i < BM; ++i) {
        const double2 v = __builtin_bit_cast(double2, As[cur][wm * (TM / 2) + i * 16 + lr][4 * q + u]);
        ar[i] = v.x;
        ai[i] = v.y;
        a3[i] = v.x + v.y;
      }
#pragma unroll
      for (int j = 0; j < BN; ++j) {
        const double2 v = __builtin_bit_cast(double2, Bs[cur][wn * (TN / 2) + j * 16 + lr][4 * q + u]);
        br[j] = v.x;
        bi[j] = v.y;
        b3[j] = v.x + v.y;
      }
#pragma unroll
      for (int i = 0; i < BM; ++i)
#pragma unroll
        for (int j = 0; j < BN; ++j) {
          acc[0][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], acc[0][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], acc[1][i][j], 0, 0, 0);
          acc[2][i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a3[i], b3[j], acc[2][i][j], 0, 0, 0);
        }
    }
    __builtin_amdgcn_s_setprio(0);
    if (t + 1 < nst) store(cur ^ 1);
    __syncthreads();
  }

  double2* Cb = p.Cd + z0 * p.sC0 + z1 * p.sC1;
#pragma unroll
  for (int i = 0; i < BM; ++i)
#pragma unroll
    for (int j = 0; j < BN; ++j) {
      const int n = n0 + wn * (TN / 2) + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * (TM / 2) + i * 16 + q + 4 * r;
        if (m < p.M && n < p.N) {
          const double p1 = acc[0][i][j][r], p2 = acc[1][i][j][r], p3 = acc[2][i][j][r];
          Cb[m * p.sCm + n * p.sCn] = double2{p1 - p2, p3 - (p1 + p2)};
        }
      }
    }
}

template <int BM, int BN>
hipError_t pk64_launch(const LpDesc& p_in, hipStream_t s) {
  const int gm = (p_in.M + 32 * BM - 1) / (32 * BM), gn = (p_in.N + 32 * BN - 1) / (32 * BN);
  const int64_t total = (int64_t)gm * gn * p_in.NB;
  if (total > 0x7fffffff) return hipErrorInvalidConfiguration;
  // Row-tiles per group: tall grids (the first stage: L's rows against x) walk 16 row-tiles per
  // column panel, square ones (the last stage) 2; measured FETCH_SIZE at D = 512, w = 3:
  // 67 and 67 MB per launch against 76 and 76 MB with 8 for both.  Timing is unchanged.
  LpDesc p = p_in;
  if (p.group <= 0) p.group = gm >= 4 * gn ? 16 : 2;
  hipLaunchKernelGGL((pk64_brgemm_kernel<BM, BN>), dim3((unsigned)total), dim3(kT), 0, s, p, gm, gn);
  return hipGetLastError();
}

template <int PREC, int BM, int BN>
hipError_t lp_launch(const LpDesc& p_in, hipStream_t s) {
  const int gm = (p_in.M + 32 * BM - 1) / (32 * BM), gn = (p_in.N + 32 * BN - 1) / (32 * BN);
  const int64_t total = (int64_t)gm * gn * p_in.NB;
  if (total > 0x7fffffff) return hipErrorInvalidConfiguration;
  // row-tiles per group (measured FETCH_SIZE at D = 512, w = 3, 64 x 32 tiles: the first stage
  // 33.6 MB with 8 against 37.8-42.2 with 2, 4 or 16; the last stage 35.7 MB with 2 against 48.3
  // with 8); timing is unchanged
  LpDesc p = p_in;
  if (p.group <= 0) p.group = gm >= gn ? kGroupM : 2;
  hipLaunchKernelGGL((lp_brgemm_kernel<PREC, BM, BN>), dim3((unsigned)total), dim3(kT), 0, s, p, gm, gn);
  return hipGetLastError();
}

template <int PREC>
hipError_t lp_launch_tile(const LpDesc& p, int tile, hipStream_t s) {
  switch (tile) {
    case 2: return lp_launch<PREC, 2, 1>(p, s);
    case 3: return lp_launch<PREC, 1, 2>(p, s);
    case 4: return lp_launch<PREC, 2, 2>(p, s);
    default: return lp_launch<PREC, 1, 1>(p, s);
  }
}

unsigned grid_for(int64_t total) { return (unsigned)std::min<int64_t>(8192, std::max<int64_t>(1, (total + kT - 1) / kT)); }

}  // namespace

int lp_auto_tile(const LpDesc& p) {
  // the largest tile that still gives two workgroups per CU.  FP64 stops at two blocks per wave
  // (the 2 x 2 form holds two waves per SIMD, too few for the FP64 MFMA) and prefers 32 x 64
  // (measured at D = 512: stage 1 87.8 us against 91.7 us for 64 x 32)
  auto wgs = [&](int bm, int bn) {
    return (int64_t)((p.M + 32 * bm - 1) / (32 * bm)) * ((p.N + 32 * bn - 1) / (32 * bn)) * p.NB;
  };
  if (p.prec == MPSD_PREC_FP64) return wgs(1, 2) >= 512 ? 3 : wgs(2, 1) >= 512 ? 2 : 1;
  if (wgs(2, 2) >= 512) return 4;
  if (p.M >= p.N && wgs(2, 1) >= 512) return 2;
  if (wgs(1, 2) >= 512) return 3;
  if (wgs(2, 1) >= 512) return 2;
  return 1;
}

hipError_t launch_lp_brgemm(const LpDesc& p, int tile, hipStream_t s) {
  if (p.M <= 0 || p.N <= 0 || p.NB <= 0) return hipSuccess;
  if (p.Kp % lp_stage(p.prec) != 0 || p.zdiv <= 0 || p.NB % p.zdiv != 0) return hipErrorInvalidValue;
  if (tile <= 0) tile = lp_auto_tile(p);
  if (p.prec == MPSD_PREC_FP64) {
    switch (tile) {
      case 2: return pk64_launch<2, 1>(p, s);
      case 3: return pk64_launch<1, 2>(p, s);
      case 4: return pk64_launch<2, 2>(p, s);
      default: return pk64_launch<1, 1>(p, s);
    }
  }
  if (p.prec == MPSD_PREC_BF16X3) return lp_launch_tile<MPSD_PREC_BF16X3>(p, tile, s);
  if (p.prec == MPSD_PREC_FP32) return lp_launch_tile<MPSD_PREC_FP32>(p, tile, s);
  return hipErrorInvalidValue;
}

hipError_t launch_lp_pack(const LpPackDesc& p, int prec, hipStream_t s) {
  const int64_t total = (int64_t)p.nb * p.rows * (p.Kp / (prec == MPSD_PREC_FP64 ? 4 : 8));
  if (total <= 0) return hipSuccess;
  if (p.Kp % lp_stage(prec) != 0) return hipErrorInvalidValue;
  if (prec == MPSD_PREC_FP64)
    hipLaunchKernelGGL(pk64_pack_kernel, dim3(grid_for(total)), dim3(kT), 0, s, p);
  else if (prec == MPSD_PREC_BF16X3)
    hipLaunchKernelGGL(lp_pack_kernel<MPSD_PREC_BF16X3>, dim3(grid_for(2 * total)), dim3(kT), 0, s, p);
  else if (prec == MPSD_PREC_FP32)
    hipLaunchKernelGGL(lp_pack_kernel<MPSD_PREC_FP32>, dim3(grid_for(2 * total)), dim3(kT), 0, s, p);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_lp_mix(const float2* T1, void* T2p, const double2* mix, int Db, int nin, int nout, int Dkr,
                         int Kp, int prec, hipStream_t s) {
  const int64_t total = (int64_t)Db * (Kp / 8) * ((nout + 1) / 2);
  if (total <= 0) return hipSuccess;
  if (nin * nout > 256 || Kp % 32 != 0) return hipErrorInvalidValue;
  // grids under kLpMixSplitBelow workgroups (D <= 256 at w = 3) run the split form: measured
  // 10-18% faster matvecs at D = 64-128, slower ones at D = 400-512 (its loads repeat per output)
  if (grid_for(total) < kLpMixSplitBelow) {
    const int64_t total_split = (int64_t)Db * (Kp / 8) * 2;
    if (prec == MPSD_PREC_BF16X3)
      hipLaunchKernelGGL(lp_mix_split_kernel<MPSD_PREC_BF16X3>, dim3(grid_for(total_split), nout), dim3(kT), 0, s, T1,
                         T2p, mix, Db, nin, nout, Dkr, Kp);
    else if (prec == MPSD_PREC_FP32)
      hipLaunchKernelGGL(lp_mix_split_kernel<MPSD_PREC_FP32>, dim3(grid_for(total_split), nout), dim3(kT), 0, s, T1,
                         T2p, mix, Db, nin, nout, Dkr, Kp);
    else
      return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (prec == MPSD_PREC_BF16X3)
    hipLaunchKernelGGL(lp_mix_kernel<MPSD_PREC_BF16X3>, dim3(grid_for(total)), dim3(kT), 0, s, T1, T2p, mix, Db, nin,
                       nout, Dkr, Kp);
  else if (prec == MPSD_PREC_FP32)
    hipLaunchKernelGGL(lp_mix_kernel<MPSD_PREC_FP32>, dim3(grid_for(total)), dim3(kT), 0, s, T1, T2p, mix, Db, nin,
                       nout, Dkr, Kp);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_pk64_mix(const double2* T1, double2* T2p, const double2* mix, int Db, int nin, int nout, int Dkr,
                           int Kp, hipStream_t s) {
  const int64_t total = (int64_t)Db * Kp;
  if (total <= 0) return hipSuccess;
  if (Kp % 16 != 0 || nout > kMixMax) return hipErrorInvalidValue;
  const unsigned g = grid_for(total);
  if (g < kMixSplitBelow)
    hipLaunchKernelGGL(pk64_mix_split_kernel, dim3(g, nout), dim3(kT), 0, s, T1, T2p, mix, Db, nin, nout, Dkr, Kp);
  else
    hipLaunchKernelGGL(pk64_mix_kernel, dim3(g), dim3(kT), 0, s, T1, T2p, mix, Db, nin, nout, Dkr, Kp);
  return hipGetLastError();
}

hipError_t launch_lp_sum(const float2* P, int np, int64_t stride, const double2* x, double beta, double2* y,
                         int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(lp_sum_kernel, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, s, P, np, stride, x, beta, y,
                     n);
  return hipGetLastError();
}

}  // namespace mpsd
