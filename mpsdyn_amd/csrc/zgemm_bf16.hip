// Split-BF16 complex BRGEMM for the precision form of the matvec (gfx950).
//
// Each FP64 operand x is split on its way into LDS as x ~ hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (round to nearest even, via FP32).  A real product a*b is taken as
// a_hi*b_hi + a_hi*b_lo + a_lo*b_hi (the lo*lo term, ~2^-16 relative, is dropped), so a
// complex 16x16x32 block costs 12 v_mfma_f32_16x16x32_bf16 with FP32 accumulation.
// Result accuracy is ~1e-5 relative; the eigensolver around it stays FP64 and refines.
//
// Tile: 32x32 outputs per workgroup (4 wave64s, one 16x16 block each), K = 32 per LDS
// stage.  LDS holds four bf16 planes (re hi/lo, im hi/lo) per operand, rows of 32 k values
// padded to 40 so the 16-byte fragment reads of a wave hit distinct banks.
// v_mfma_f32_16x16x32_bf16 layout: A[m][k] from lane m + 16*(k/8), 8 consecutive k per
// lane; B[k][n] from lane n + 16*(k/8); result (m, n) in lane n + 16*(m/4), register m%4.

#include "mpsd_internal.h"

namespace mpsd {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int TM = 32, TN = 32, TK = 32, PITCH = 40;
constexpr int kGroupM = 8;

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ void split(double x, unsigned short& hi, unsigned short& lo) {
  const float f = (float)x;
  hi = f2bf(f);
  lo = f2bf(f - bf2f(hi));
}

__device__ __forceinline__ void tile_of(int gm, int gn, int& tm_i, int& tn_i, int& z) {
  const int total = gridDim.x;
  int id = blockIdx.x;
  if ((total & 7) == 0) id = (id & 7) * (total >> 3) + (id >> 3);
  const int per_z = gm * gn;
  z = id / per_z;
  const int r = id - z * per_z;
  const int in_group = kGroupM * gn;
  const int g = r / in_group;
  const int first_m = g * kGroupM;
  const int gsize = min(gm - first_m, kGroupM);
  const int rr = r - g * in_group;
  tm_i = first_m + rr % gsize;
  tn_i = rr / gsize;
}

// One operand tile in LDS: [plane][row][k], planes = re hi, re lo, im hi, im lo.
struct alignas(16) Planes {
  unsigned short v[4][32][PITCH];
};

__device__ __forceinline__ bf16x8 frag(const Planes& P, int plane, int row, int k0) {
  const u4v u = *reinterpret_cast<const u4v*>(&P.v[plane][row][k0]);
  return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ bf16x8 neg(bf16x8 a) {
  u4v u = __builtin_bit_cast(u4v, a);
  u ^= u4v{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ f4 mma(bf16x8 a, bf16x8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

template <int NBW, bool ASH, bool BSH>
__global__ void __launch_bounds__(kThreads) zbrgemm_bf16x3_kernel(const ZgemmDesc p, int gm, int gn) {
  constexpr int NA = ASH ? 1 : NBW;
  constexpr int NB_ = BSH ? 1 : NBW;
  constexpr int AE = TM * TK / kThreads;  // 4
  constexpr int BE = TK * TN / kThreads;  // 4
  __shared__ Planes As[NA];
  __shared__ Planes Bs[NB_];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  int tm_i, tn_i, z;
  tile_of(gm, gn, tm_i, tn_i, z);
  const int m0 = tm_i * TM, n0 = tn_i * TN;
  const int nbw = min(NBW, p.NB - z * NBW);

  f4 accr[NBW], acci[NBW];
#pragma unroll
  for (int j = 0; j < NBW; ++j) accr[j] = acci[j] = f4{0.f, 0.f, 0.f, 0.f};

  double2 ra[NA][AE], rb[NB_][BE];
  const int nk = (p.K + TK - 1) / TK;
  const int ntiles = nk * p.R;

  auto a_coord = [&](int e, int& m, int& k) {
    const int idx = tid + kThreads * e;
    if (p.a_mfast) {
      m = idx % TM;
      k = idx / TM;
    } else {
      k = idx % TK;
      m = idx / TK;
    }
  };
  auto b_coord = [&](int e, int& k, int& n) {
    const int idx = tid + kThreads * e;
    if (p.b_kfast) {
      k = idx % TK;
      n = idx / TK;
    } else {
      n = idx % TN;
      k = idx / TN;
    }
  };
  auto load_tile = [&](int t) {
    const int r = t / nk;
    const int k0 = (t - r * nk) * TK;
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int jb = (j < nbw) ? j : 0;
      const double2* base = p.A + (int64_t)z * p.sAz + (int64_t)jb * p.sAb + (int64_t)r * p.sAr;
#pragma unroll
      for (int e = 0; e < AE; ++e) {
        int m, k;
        a_coord(e, m, k);
        const int gmm = m0 + m, gk = k0 + k;
        ra[j][e] = (gmm < p.M && gk < p.K) ? base[gmm * p.sAm + gk * p.sAk] : double2{0.0, 0.0};
      }
    }
#pragma unroll
    for (int j = 0; j < NB_; ++j) {
      const int jb = (j < nbw) ? j : 0;
      const double2* base = p.B + (int64_t)z * p.sBz + (int64_t)jb * p.sBb + (int64_t)r * p.sBr;
#pragma unroll
      for (int e = 0; e < BE; ++e) {
        int k, n;
        b_coord(e, k, n);
        const int gk = k0 + k, gnn = n0 + n;
        rb[j][e] = (gk < p.K && gnn < p.N) ? base[gk * p.sBk + gnn * p.sBn] : double2{0.0, 0.0};
      }
    }
  };
  auto store_tile = [&]() {
    const double csign = p.conjA ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < NA; ++j)
#pragma unroll
      for (int e = 0; e < AE; ++e) {
        int m, k;
        a_coord(e, m, k);
        unsigned short h, l;
        split(ra[j][e].x, h, l);
        As[j].v[0][m][k] = h;
        As[j].v[1][m][k] = l;
        split(csign * ra[j][e].y, h, l);
        As[j].v[2][m][k] = h;
        As[j].v[3][m][k] = l;
      }
#pragma unroll
    for (int j = 0; j < NB_; ++j)
#pragma unroll
      for (int e = 0; e < BE; ++e) {
        int k, n;
        b_coord(e, k, n);
        unsigned short h, l;
        split(rb[j][e].x, h, l);
        Bs[j].v[0][n][k] = h;
        Bs[j].v[1][n][k] = l;
        split(rb[j][e].y, h, l);
        Bs[j].v[2][n][k] = h;
        Bs[j].v[3][n][k] = l;
      }
  };

  const int frow_a = wm * 16 + (lane & 15);
  const int frow_b = wn * 16 + (lane & 15);
  const int fk = 8 * (lane >> 4);
  if (ntiles > 0) load_tile(0);
  for (int t = 0; t < ntiles; ++t) {
    store_tile();
    __syncthreads();
    if (t + 1 < ntiles) load_tile(t + 1);
    bf16x8 a[NA][4], b[NB_][4];
#pragma unroll
    for (int j = 0; j < NA; ++j)
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) a[j][pl] = frag(As[j], pl, frow_a, fk);
#pragma unroll
    for (int j = 0; j < NB_; ++j)
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) b[j][pl] = frag(Bs[j], pl, frow_b, fk);
#pragma unroll
    for (int j = 0; j < NBW; ++j) {
      const int ja = ASH ? 0 : j, jb = BSH ? 0 : j;
      const bf16x8 nih = neg(a[ja][2]), nil = neg(a[ja][3]);
      f4 cr = accr[j], ci = acci[j];
      cr = mma(a[ja][0], b[jb][0], cr);
      ci = mma(a[ja][0], b[jb][2], ci);
      cr = mma(a[ja][0], b[jb][1], cr);
      ci = mma(a[ja][0], b[jb][3], ci);
      cr = mma(a[ja][1], b[jb][0], cr);
      ci = mma(a[ja][1], b[jb][2], ci);
      cr = mma(nih, b[jb][2], cr);
      ci = mma(a[ja][2], b[jb][0], ci);
      cr = mma(nih, b[jb][3], cr);
      ci = mma(a[ja][2], b[jb][1], ci);
      cr = mma(nil, b[jb][2], cr);
      ci = mma(a[ja][3], b[jb][0], ci);
      accr[j] = cr;
      acci[j] = ci;
    }
    __syncthreads();
  }

  const int gn_ = n0 + wn * 16 + (lane & 15);
#pragma unroll
  for (int j = 0; j < NBW; ++j) {
    if (j >= nbw) break;
    double2* cb = p.C + (int64_t)z * p.sCz + (int64_t)j * p.sCb;
    const double2* db = p.Dadd ? p.Dadd + (int64_t)z * p.sCz + (int64_t)j * p.sCb : nullptr;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int gmm = m0 + wm * 16 + 4 * (lane >> 4) + rr;
      if (gmm < p.M && gn_ < p.N) {
        const int64_t off = gmm * p.sCm + gn_ * p.sCn;
        double2 v{(double)accr[j][rr], (double)acci[j][rr]};
        if (db) {
          const double2 x = db[off];
          v.x += p.beta * x.x;
          v.y += p.beta * x.y;
        }
        cb[off] = v;
      }
    }
  }
}

template <int NBW, bool ASH, bool BSH>
hipError_t launch(const ZgemmDesc& p, hipStream_t s) {
  const int gm = (p.M + TM - 1) / TM, gn = (p.N + TN - 1) / TN, gz = (p.NB + NBW - 1) / NBW;
  const int64_t total = (int64_t)gm * gn * gz;
  if (total > 0x7fffffff) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((zbrgemm_bf16x3_kernel<NBW, ASH, BSH>), dim3((unsigned)total), dim3(kThreads), 0, s, p, gm, gn);
  return hipGetLastError();
}

}  // namespace

// p must already have its group strides resolved (see launch_zbrgemm_f64).
hipError_t launch_zbrgemm_bf16x3(const ZgemmDesc& p, int nbw, hipStream_t s) {
  if (nbw == 2 && p.sAb == 0) return launch<2, true, false>(p, s);
  if (nbw == 2 && p.sBb == 0) return launch<2, false, true>(p, s);
  if (nbw == 2) return launch<2, false, false>(p, s);
  return launch<1, false, false>(p, s);
}

}  // namespace mpsd
