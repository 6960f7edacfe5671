// FP64 complex contraction kernels for gfx950 (CDNA4).
//
// zbrgemm_f64: batch-reduce complex GEMM on v_mfma_f64_16x16x4f64.  A workgroup of four
// 64-wide waves owns a TM x TN output tile for NBW batches; operand tiles are staged
// through LDS as interleaved (re, im) pairs.  A complex 16x16x4 block costs three real
// MFMAs in FP64 (the three-product form: re*re, im*im and (re+im)*(re+im), combined in the
// epilogue) and four in FP32 (re*re, -im*im into the real accumulator; re*im, im*re into
// the imaginary one).
// The global->register load of tile t+1 is issued before the MFMAs of tile t, so HBM/L2
// latency hides behind the matrix cores.  When one operand is shared by all batches of
// the workgroup (the environment in stage 1, the right environment in stage 3) it is
// staged once and reused NBW times.
//
// MFMA 16x16x4 f64 lane layout (wave64): A[m][k] comes from lane m + 16k, B[k][n] from
// lane n + 16k, and result element (m, n) lands in lane n + 16*(m%4), register m/4
// (the f64 form interleaves rows across the four lane groups, unlike the f32 forms).

#include "mpsd_internal.h"

namespace mpsd {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kThreads = 256;
// Threads of a BRGEMM workgroup: 2 x 2 waves, each with a (TM/2) x (TN/2) block of the tile;
// a 16 x 16 tile is one wave.
constexpr int tile_threads(int tm, int tn) { return tm * tn <= 256 ? 64 : kThreads; }

// Arithmetic of the BRGEMM's MFMAs.  Operands always arrive as FP64 from HBM and leave as
// FP64; the FP32 form converts when staging into LDS and accumulates in FP32.
template <int ARITH>
struct Pol;
template <>
struct Pol<0> {  // FP64: v_mfma_f64_16x16x4f64, element (m, n) in lane n + 16*(m%4), reg m/4
  typedef double T;
  typedef double2 T2;
  typedef d4 Acc;
  static __device__ __forceinline__ Acc mfma(T a, T b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
  static __device__ __forceinline__ T2 cvt(double2 v, double s) { return T2{v.x, s * v.y}; }
};
template <>
struct Pol<1> {  // FP32: v_mfma_f32_16x16x4f32, element (m, n) in lane n + 16*(m/4), reg m%4
  typedef float T;
  typedef float2 T2;
  typedef f4 Acc;
  static __device__ __forceinline__ Acc mfma(T a, T b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  static __device__ __forceinline__ T2 cvt(double2 v, double s) { return T2{(float)v.x, (float)(s * v.y)}; }
};

__device__ __forceinline__ double2 cload(const double2* p) { return *p; }

// Workgroup -> tile mapping.  The hardware hands linear workgroup ids to the 8 XCDs round
// robin; the ids are first regrouped so each XCD owns a contiguous range.  z (the batch
// group) is slowest.  When the tile grid of a z splits evenly into px x py = 8 blocks, the
// ranges are those blocks, with (px, py) chosen to minimise the row and column tiles each XCD
// touches: one XCD's L2 then fetches (gm/px) A panels and (gn/py) B panels instead of a
// strip of A rows and nearly all of B (the D = 512 env x site stage: 4 x 2 blocks fetch
// about 57 MB instead of 77 MB).  Each block (or the whole grid) is walked in groups of
// kGroupM row-tiles so the tiles resident together share operand panels.  The mapping only
// decides which workgroup computes a tile, never the order of its sums.
constexpr int kGroupM = 8;

__device__ __forceinline__ void tile_of(int gm, int gn, int& tm_i, int& tn_i, int& z) {
  const int total = gridDim.x;
  int id = blockIdx.x;
  if ((total & 7) == 0) id = (id & 7) * (total >> 3) + (id >> 3);
  const int per_z = gm * gn;
  z = id / per_z;
  int r = id - z * per_z;
  int px = 0, py = 0;
  if ((total & 7) == 0 && (per_z & 7) == 0) {
    int best = 1 << 30;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int qx = 8 >> f, qy = 1 << f;
      if (gm % qx == 0 && gn % qy == 0 && gm / qx + gn / qy < best) {
        best = gm / qx + gn / qy;
        px = qx;
        py = qy;
      }
    }
  }
  int bm = gm, bn = gn, om = 0, on = 0;
  if (px) {
    bm = gm / px;
    bn = gn / py;
    const int blk = r / (bm * bn);
    r -= blk * (bm * bn);
    om = (blk % px) * bm;
    on = (blk / px) * bn;
  }
  const int in_group = kGroupM * bn;
  const int g = r / in_group;
  const int first_m = g * kGroupM;
  const int gsize = min(bm - first_m, kGroupM);
  const int rr = r - g * in_group;
  tm_i = om + first_m + rr % gsize;
  tn_i = on + rr / gsize;
}

// LDS operand tiles of one workgroup, [k][m] and [k][n].  Each row carries one element of
// padding: when the operand is k-fast in HBM, consecutive threads store consecutive k, and
// an unpadded 32-element row (512 B) would put all of them in the same banks.
// Tiles whose single stage fits in 40 KB are double-buffered (one barrier per K stage: the
// next stage is stored into the other buffer while this one is read).
template <int TM, int TN, int TK, int NBW, bool ASH, bool BSH, int ARITH>
struct Smem {
  typedef typename Pol<ARITH>::T2 T2;
  static constexpr int kStage = ((ASH ? 1 : NBW) * TM + (BSH ? 1 : NBW) * TN + 2) * TK * (int)sizeof(T2);
  static constexpr int NBUF = kStage <= 40960 ? 2 : 1;
  T2 As[NBUF][ASH ? 1 : NBW][TK][TM + 1];
  T2 Bs[NBUF][BSH ? 1 : NBW][TK][TN + 1];
};

// One TM x TN output tile (for NBW batches of group z) of the BRGEMM described by p.  The
// order of every accumulation depends only on p and the tile, never on which workgroup or
// launch computes it, so a problem gives bit-identical results alone and inside a group.
template <int TM, int TN, int TK, int NBW, bool ASH, bool BSH, int ARITH>
__device__ __forceinline__ void brgemm_tile(const ZgemmDesc& p, int m0, int n0, int z,
                                            Smem<TM, TN, TK, NBW, ASH, BSH, ARITH>& sm) {
  typedef Pol<ARITH> P;
  typedef typename P::T T;
  typedef typename P::T2 T2;
  typedef typename P::Acc Acc;
  constexpr int NA = ASH ? 1 : NBW;
  constexpr int NB_ = BSH ? 1 : NBW;
  constexpr int NBUF = Smem<TM, TN, TK, NBW, ASH, BSH, ARITH>::NBUF;

  constexpr int NT = tile_threads(TM, TN);
  constexpr int WG = NT / 64 == 4 ? 2 : 1;   // waves along m and along n
  constexpr int WTM = TM / WG, WTN = TN / WG;
  constexpr int BM = WTM / 16, BN = WTN / 16;
  constexpr int AE = TM * TK / NT;
  constexpr int BE = TK * TN / NT;
  static_assert(AE >= 1 && BE >= 1 && BM >= 1 && BN >= 1, "tile too small for its threads");

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = WG == 2 ? (wave & 1) : 0, wn = WG == 2 ? (wave >> 1) : 0;
  const int nbw = min(NBW, p.NB - z * NBW);

  // FP64 takes the three-product form of the complex product (P1 = ar*br, P2 = ai*bi,
  // P3 = (ar+ai)*(br+bi); re = P1 - P2, im = P3 - P1 - P2): three MFMA chains per block
  // instead of four.  FP32 keeps the four-product form (no cancellation in im).
  constexpr bool G3 = (ARITH == 0);
  constexpr int NP = G3 ? 3 : 2;
  Acc acc[NP][NBW][BM][BN];
#pragma unroll
  for (int c = 0; c < NP; ++c)
#pragma unroll
    for (int j = 0; j < NBW; ++j)
#pragma unroll
      for (int i = 0; i < BM; ++i)
#pragma unroll
        for (int q = 0; q < BN; ++q) acc[c][j][i][q] = Acc{0, 0, 0, 0};

  double2 ra[NA][AE];
  double2 rb[NB_][BE];

  const int nk = (p.K + TK - 1) / TK;
  const int ntiles = nk * p.R;

  auto a_coord = [&](int e, int& m, int& k) {
    const int idx = tid + NT * e;
    if (p.a_mfast) {
      m = idx % TM;
      k = idx / TM;
    } else {
      k = idx % TK;
      m = idx / TK;
    }
  };
  auto b_coord = [&](int e, int& k, int& n) {
    const int idx = tid + NT * e;
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
        const int gm = m0 + m, gk = k0 + k;
        ra[j][e] = (gm < p.M && gk < p.K) ? cload(base + gm * p.sAm + gk * p.sAk)
                                          : double2{0.0, 0.0};
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
        const int gk = k0 + k, gn = n0 + n;
        rb[j][e] = (gk < p.K && gn < p.N) ? cload(base + gk * p.sBk + gn * p.sBn)
                                          : double2{0.0, 0.0};
      }
    }
  };

  auto store_tile = [&](int buf) {
    auto& As = sm.As[buf];
    auto& Bs = sm.Bs[buf];
    const double csign = p.conjA ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < NA; ++j)
#pragma unroll
      for (int e = 0; e < AE; ++e) {
        int m, k;
        a_coord(e, m, k);
        As[j][k][m] = P::cvt(ra[j][e], csign);
      }
#pragma unroll
    for (int j = 0; j < NB_; ++j)
#pragma unroll
      for (int e = 0; e < BE; ++e) {
        int k, n;
        b_coord(e, k, n);
        Bs[j][k][n] = P::cvt(rb[j][e], 1.0);
      }
  };

  if (NBUF == 2 && ntiles > 0) {
    load_tile(0);
    store_tile(0);
    __syncthreads();
  } else if (ntiles > 0) {
    load_tile(0);
  }
  for (int t = 0; t < ntiles; ++t) {
    const int cur = NBUF == 2 ? (t & 1) : 0;
    if constexpr (NBUF == 1) {
      store_tile(0);
      __syncthreads();
    }
    if (t + 1 < ntiles) load_tile(t + 1);
    auto& As = sm.As[cur];
    auto& Bs = sm.Bs[cur];
    // the MFMA block at raised wave priority: the co-resident workgroups' loads and LDS stores
    // wait behind it (measured +0.4% at D = 512)
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int kk = 0; kk < TK / 4; ++kk) {
      const int krow = kk * 4 + (lane >> 4);
      const int lc = lane & 15;
      // a3/b3: re + im for the three-product form, -im for the four-product form
      T ar[NA][BM], ai[NA][BM], a3[NA][BM];
      T br[NB_][BN], bi[NB_][BN], b3[NB_][BN];
#pragma unroll
      for (int j = 0; j < NA; ++j)
#pragma unroll
        for (int i = 0; i < BM; ++i) {
          const T2 v = As[j][krow][wm * WTM + i * 16 + lc];
          ar[j][i] = v.x;
          ai[j][i] = v.y;
          a3[j][i] = G3 ? v.x + v.y : -v.y;
        }
#pragma unroll
      for (int j = 0; j < NB_; ++j)
#pragma unroll
        for (int q = 0; q < BN; ++q) {
          const T2 v = Bs[j][krow][wn * WTN + q * 16 + lc];
          br[j][q] = v.x;
          bi[j][q] = v.y;
          if constexpr (G3) b3[j][q] = v.x + v.y;
        }
#pragma unroll
      for (int j = 0; j < NBW; ++j) {
        const int ja = ASH ? 0 : j;
        const int jb = BSH ? 0 : j;
#pragma unroll
        for (int i = 0; i < BM; ++i)
#pragma unroll
          for (int q = 0; q < BN; ++q) {
            if constexpr (G3) {
              acc[0][j][i][q] = P::mfma(ar[ja][i], br[jb][q], acc[0][j][i][q]);
              acc[1][j][i][q] = P::mfma(ai[ja][i], bi[jb][q], acc[1][j][i][q]);
              acc[2][j][i][q] = P::mfma(a3[ja][i], b3[jb][q], acc[2][j][i][q]);
            } else {
              acc[0][j][i][q] = P::mfma(ar[ja][i], br[jb][q], acc[0][j][i][q]);
              acc[1][j][i][q] = P::mfma(ar[ja][i], bi[jb][q], acc[1][j][i][q]);
            }
          }
        if constexpr (!G3) {
#pragma unroll
          for (int i = 0; i < BM; ++i)
#pragma unroll
            for (int q = 0; q < BN; ++q) {
              acc[0][j][i][q] = P::mfma(a3[ja][i], bi[jb][q], acc[0][j][i][q]);
              acc[1][j][i][q] = P::mfma(ai[ja][i], br[jb][q], acc[1][j][i][q]);
            }
        }
      }
    }
    // double buffer: stage t+1 goes into the other buffer; the barrier makes it visible and
    // retires every read of this one before it is overwritten two stages later
    __builtin_amdgcn_s_setprio(0);
    if (NBUF == 2 && t + 1 < ntiles) store_tile(cur ^ 1);
    __syncthreads();
  }

  // Epilogue: C = acc (+ beta * Dadd).
  const int lc = lane & 15;
#pragma unroll
  for (int j = 0; j < NBW; ++j) {
    if (j >= nbw) break;
    double2* cb = p.C + (int64_t)z * p.sCz + (int64_t)j * p.sCb;
    const double2* db = p.Dadd ? p.Dadd + (int64_t)z * p.sCz + (int64_t)j * p.sCb : nullptr;
#pragma unroll
    for (int i = 0; i < BM; ++i)
#pragma unroll
      for (int q = 0; q < BN; ++q) {
        const int gn = n0 + wn * WTN + q * 16 + lc;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int gm = m0 + wm * WTM + i * 16 + P::row(lane, rr);
          if (gm < p.M && gn < p.N) {
            const int64_t off = gm * p.sCm + gn * p.sCn;
            double2 v;
            if constexpr (G3) {
              const double p1 = acc[0][j][i][q][rr], p2 = acc[1][j][i][q][rr], p3 = acc[2][j][i][q][rr];
              v = double2{p1 - p2, p3 - (p1 + p2)};
            } else {
              v = double2{(double)acc[0][j][i][q][rr], (double)acc[1][j][i][q][rr]};
            }
            if (db) {
              const double2 x = db[off];
              v.x = fma(p.beta, x.x, v.x);
              v.y = fma(p.beta, x.y, v.y);
            }
            cb[off] = v;
          }
        }
      }
  }
}

// Register budget of the 32x32x16 tile: FP64 MFMAs of one wave issue far apart, so the
// matrix core needs several resident waves per SIMD.  Two batches per workgroup (768 tiles
// on 256 CUs for the D=512 stages) are held to three waves per SIMD, one batch per
// workgroup (twice the tiles) to six.  Larger tiles are left to the compiler.
constexpr int min_waves(int tm, int tn, int tk, int nbw) {
  return (tm == 32 && tn == 32 && tk == 16) ? (nbw == 1 ? 6 : 3) : 1;
}

template <int TM, int TN, int TK, int NBW, bool ASH, bool BSH, int ARITH>
__global__ void __launch_bounds__(tile_threads(TM, TN)) __attribute__((amdgpu_waves_per_eu(min_waves(TM, TN, TK, NBW)))) zbrgemm_kernel(const ZgemmDesc p, int gm, int gn) {
  __shared__ Smem<TM, TN, TK, NBW, ASH, BSH, ARITH> sm;
  int tm_i, tn_i, z;
  tile_of(gm, gn, tm_i, tn_i, z);
  brgemm_tile<TM, TN, TK, NBW, ASH, BSH, ARITH>(p, tm_i * TM, tn_i * TN, z, sm);
}

// Grouped persistent form: the workgroups of one launch stride over the concatenated tile
// lists of all problems (problem q owns tiles [tile_begin[q], tile_begin[q+1]) ).
template <int TM, int TN, int TK, int NBW, bool ASH, bool BSH, int ARITH>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(min_waves(TM, TN, TK, NBW)))) zbrgemm_grouped_kernel(const GroupedProblem* __restrict__ probs, int nprob,
                                                                   const int* __restrict__ nprob_dev) {
  __shared__ Smem<TM, TN, TK, NBW, ASH, BSH, ARITH> sm;
  if (nprob_dev) nprob = *nprob_dev;
  if (nprob <= 0) return;
  const int total = probs[nprob - 1].tile_begin + probs[nprob - 1].ntiles;
  int q = 0;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    while (t >= probs[q].tile_begin + probs[q].ntiles) ++q;  // tiles are visited in increasing order
    const GroupedProblem& gp = probs[q];
    const int local = t - gp.tile_begin;
    const int per_z = gp.gm * gp.gn;
    const int z = local / per_z;
    const int r = local - z * per_z;
    brgemm_tile<TM, TN, TK, NBW, ASH, BSH, ARITH>(gp.desc, (r % gp.gm) * TM, (r / gp.gm) * TN, z, sm);
  }
}

template <int TM, int TN, int TK, int NBW, bool ASH, bool BSH, int ARITH>
hipError_t launch_cfg_a(const ZgemmDesc& p, hipStream_t s) {
  const int gm = (p.M + TM - 1) / TM, gn = (p.N + TN - 1) / TN, gz = (p.NB + NBW - 1) / NBW;
  const int64_t total = (int64_t)gm * gn * gz;
  if (total > 0x7fffffff) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((zbrgemm_kernel<TM, TN, TK, NBW, ASH, BSH, ARITH>), dim3((unsigned)total),
                     dim3(tile_threads(TM, TN)), 0,
                     s, p, gm, gn);
  return hipGetLastError();
}

template <int TM, int TN, int TK, int NBW, bool ASH, bool BSH>
hipError_t launch_cfg(const ZgemmDesc& p, hipStream_t s) {
  if (p.arith == 1) return launch_cfg_a<TM, TN, TK, NBW, ASH, BSH, 1>(p, s);
  return launch_cfg_a<TM, TN, TK, NBW, ASH, BSH, 0>(p, s);
}

template <int TM, int TN, int TK>
hipError_t launch_tile(const ZgemmDesc& p, int nbw, hipStream_t s) {
  if (nbw == 2 && p.sAb == 0) return launch_cfg<TM, TN, TK, 2, true, false>(p, s);
  if (nbw == 2 && p.sBb == 0) return launch_cfg<TM, TN, TK, 2, false, true>(p, s);
  if (nbw == 2) return launch_cfg<TM, TN, TK, 2, false, false>(p, s);
  return launch_cfg<TM, TN, TK, 1, false, false>(p, s);
}

template <int TK>
hipError_t launch_tk(const ZgemmDesc& p, TileCfg cfg, int nbw, hipStream_t s) {
  if (cfg.tm == 16 && cfg.tn == 16) return launch_tile<16, 16, TK>(p, nbw, s);
  if (cfg.tm == 64 && cfg.tn == 64) return launch_tile<64, 64, TK>(p, nbw, s);
  if (cfg.tm == 64 && cfg.tn == 32) return launch_tile<64, 32, TK>(p, nbw, s);
  if (cfg.tm == 32 && cfg.tn == 64) return launch_tile<32, 64, TK>(p, nbw, s);
  return launch_tile<32, 32, TK>(p, nbw, s);
}

}  // namespace

int choose_group(ZgemmDesc& p, int cfg_nbw) {
  // One batch per workgroup on request (twice the workgroups, six waves per SIMD for the
  // FP64 32x32x16 tile); by default two batches share the common operand's LDS tile.  The
  // per-element sums are the same either way.
  if (p.group == 0 && p.NB >= 2 && cfg_nbw == 1) {
    p.group = 1;
    p.sAz = p.sAb;
    p.sBz = p.sBb;
    p.sCz = p.sCb;
  }
  return resolve_group(p);
}

int resolve_group(ZgemmDesc& p) {
  // batches per workgroup: an explicit group, else two when one operand is shared by all
  if (p.group > 0) return (p.group > 2 || p.NB % p.group != 0) ? -1 : p.group;
  const int nbw = (p.NB >= 2 && (p.sAb == 0 || p.sBb == 0)) ? 2 : 1;
  p.sAz = nbw * p.sAb;
  p.sBz = nbw * p.sBb;
  p.sCz = nbw * p.sCb;
  p.group = nbw;
  return nbw;
}

namespace {
template <int TM, int TN, int TK, int NBW, bool ASH, bool BSH>
hipError_t grouped_a(const GroupedProblem* d, int n, const int* nd, int grid, int arith, hipStream_t s) {
  if (arith == 1)
    hipLaunchKernelGGL((zbrgemm_grouped_kernel<TM, TN, TK, NBW, ASH, BSH, 1>), dim3(grid), dim3(kThreads), 0, s, d, n,
                       nd);
  else
    hipLaunchKernelGGL((zbrgemm_grouped_kernel<TM, TN, TK, NBW, ASH, BSH, 0>), dim3(grid), dim3(kThreads), 0, s, d, n,
                       nd);
  return hipGetLastError();
}
template <int TM, int TN, int TK>
hipError_t grouped_k(const GroupedProblem* d, int n, const int* nd, int grid, int nbw, int kind, int arith,
                     hipStream_t s) {
  if (nbw == 2 && kind == 1) return grouped_a<TM, TN, TK, 2, true, false>(d, n, nd, grid, arith, s);
  if (nbw == 2 && kind == 2) return grouped_a<TM, TN, TK, 2, false, true>(d, n, nd, grid, arith, s);
  if (nbw == 2) return grouped_a<TM, TN, TK, 2, false, false>(d, n, nd, grid, arith, s);
  return grouped_a<TM, TN, TK, 1, false, false>(d, n, nd, grid, arith, s);
}
}  // namespace

hipError_t launch_zbrgemm_grouped(const GroupedProblem* d, int n, int grid, int tm, int tn, int tk, int nbw, int kind,
                                  int arith, hipStream_t s, const int* nprob_dev) {
  if ((n <= 0 && !nprob_dev) || grid <= 0) return hipSuccess;
  if (arith != 0 && arith != 1) return hipErrorInvalidValue;
  if (tm == 32 && tn == 32) return tk == 32 ? grouped_k<32, 32, 32>(d, n, nprob_dev, grid, nbw, kind, arith, s)
                                            : grouped_k<32, 32, 16>(d, n, nprob_dev, grid, nbw, kind, arith, s);
  if (tm == 64 && tn == 32) return grouped_k<64, 32, 16>(d, n, nprob_dev, grid, nbw, kind, arith, s);
  if (tm == 32 && tn == 64) return grouped_k<32, 64, 16>(d, n, nprob_dev, grid, nbw, kind, arith, s);
  return hipErrorInvalidValue;
}

hipError_t launch_zbrgemm_f64(const ZgemmDesc& pin, TileCfg cfg, hipStream_t s) {
  if (pin.M <= 0 || pin.N <= 0 || pin.NB <= 0) return hipSuccess;
  ZgemmDesc p = pin;
  if (cfg.nbw == 0 && p.group == 0 && p.NB >= 2 && p.arith != 2 && (p.sAb == 0 || p.sBb == 0) &&
      (int64_t)((p.M + 31) / 32) * ((p.N + 31) / 32) * ((p.NB + 1) / 2) < 512) {
    // Two batches per workgroup share the common operand's LDS tile.  Below 512 such 32x32
    // workgroups (D <= 256 at w = 3), one batch each: twice the workgroups, half the MFMA
    // chains per wave; 8-12% faster matvecs at D = 64..256.  The sums are the same either way.
    cfg.nbw = 1;
  }
  const int nbw = choose_group(p, cfg.nbw);
  if (nbw < 0) return hipErrorInvalidValue;
  if (p.arith == 2) return launch_zbrgemm_bf16x3(p, nbw, s);  // fixed 32x32 tiles
  if ((cfg.tm == 0 || cfg.tn == 0) && p.arith == 0) {
    // FP64: 32x32 on every shape measured (D = 64..1024, w = 3, 5).  Larger FP64 tiles hold
    // one or two waves per SIMD, too few to keep the FP64 MFMA issuing.  One-wave 16x16 tiles
    // (four times the workgroups, same bits) measured 5-14% slower at D = 64..256: a wave's
    // chain of K stages is as long as in the 32x32 tile, and it loads four times the operands.
    cfg.tm = 32, cfg.tn = 32;
  } else if (cfg.tm == 0 || cfg.tn == 0) {
    // Largest tile that still gives >= 512 workgroups (two per CU); 32x32 otherwise.
    const int64_t zb = (p.NB + nbw - 1) / nbw;
    auto wgs = [&](int tm, int tn) { return (int64_t)((p.M + tm - 1) / tm) * ((p.N + tn - 1) / tn) * zb; };
    if (wgs(64, 64) >= 512)
      cfg.tm = 64, cfg.tn = 64;
    else if (wgs(64, 32) >= 512)
      cfg.tm = 64, cfg.tn = 32;
    else
      cfg.tm = 32, cfg.tn = 32;
  }
  if (cfg.tk == 32) return launch_tk<32>(p, cfg, nbw, s);
  return launch_tk<16>(p, cfg, nbw, s);
}

// ------------------------------------------------------------------------------ MPO mix

namespace {

constexpr int kMixMax = 16;

__global__ void __launch_bounds__(kThreads) zmix_f64_kernel(const double2* __restrict__ in,
                                                            double2* __restrict__ out,
                                                            const double2* __restrict__ mix,
                                                            int Bx, int nin, int nout, int ldmix,
                                                            int ldout, int64_t total) {
  const int64_t gid = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (gid >= total) return;
  const int bx = (int)(gid % Bx);
  const int64_t ax = gid / Bx;
  const double2* src = in + bx + (int64_t)Bx * nin * ax;
  double2 acc[kMixMax];
#pragma unroll
  for (int o = 0; o < kMixMax; ++o) acc[o] = double2{0.0, 0.0};
  for (int i = 0; i < nin; ++i) {
    const double2 v = src[(int64_t)Bx * i];
#pragma unroll
    for (int o = 0; o < kMixMax; ++o) {
      if (o < nout) {
        const double2 w = mix[o + ldmix * i];
        acc[o].x = fma(w.x, v.x, fma(-w.y, v.y, acc[o].x));
        acc[o].y = fma(w.x, v.y, fma(w.y, v.x, acc[o].y));
      }
    }
  }
  double2* dst = out + bx + (int64_t)Bx * ldout * ax;
#pragma unroll
  for (int o = 0; o < kMixMax; ++o)
    if (o < nout) dst[(int64_t)Bx * o] = acc[o];
}

__global__ void __launch_bounds__(kThreads) heff_full_f64_kernel(
    const double2* __restrict__ L, const double2* __restrict__ W, const double2* __restrict__ R,
    double2* __restrict__ H, int d, int Db, int wl, int Dk, int Dbr, int wr, int Dkr,
    double offset, int64_t nrows, int64_t total) {
  const int64_t gid = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (gid >= total) return;
  const int64_t row = gid % nrows, col = gid / nrows;
  const int sp = (int)(row % d);
  const int b = (int)((row / d) % Db);
  const int c = (int)(row / ((int64_t)d * Db));
  const int s = (int)(col % d);
  const int a = (int)((col / d) % Dk);
  const int a2 = (int)(col / ((int64_t)d * Dk));
  double2 acc{0.0, 0.0};
  for (int al = 0; al < wl; ++al) {
    const double2 l = L[b + (int64_t)Db * al + (int64_t)Db * wl * a];
    for (int be = 0; be < wr; ++be) {
      const double2 w = W[sp + d * al + (int64_t)d * wl * s + (int64_t)d * wl * d * be];
      const double2 r = R[c + (int64_t)Dbr * be + (int64_t)Dbr * wr * a2];
      const double2 lw{l.x * w.x - l.y * w.y, l.x * w.y + l.y * w.x};
      acc.x += lw.x * r.x - lw.y * r.y;
      acc.y += lw.x * r.y + lw.y * r.x;
    }
  }
  if (row == col) acc.x += offset;
  H[gid] = acc;
}

}  // namespace

namespace {
// y[i] = sum_{p < np} P[p * stride + i] (+ beta * x[i]), summed in a fixed order.
__global__ void __launch_bounds__(kThreads) zsum_parts_kernel(const double2* __restrict__ P, int np, int64_t stride,
                                                              const double2* __restrict__ x, double beta,
                                                              double2* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  double2 acc = P[i];
  for (int q = 1; q < np; ++q) {
    const double2 v = P[(int64_t)q * stride + i];
    acc.x += v.x;
    acc.y += v.y;
  }
  if (x) {
    const double2 v = x[i];
    acc.x = fma(beta, v.x, acc.x);
    acc.y = fma(beta, v.y, acc.y);
  }
  y[i] = acc;
}
}  // namespace

hipError_t launch_zsum_parts_f64(const double2* P, int np, int64_t stride, const double2* x, double beta, double2* y,
                                 int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zsum_parts_kernel, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, P, np,
                     stride, x, beta, y, n);
  return hipGetLastError();
}

namespace {
__device__ __forceinline__ int find_elem(const GroupedElem* probs, int nprob, int64_t i, int q) {
  while (q + 1 < nprob && i >= probs[q + 1].begin) ++q;
  return q;
}

// Persistent grid-stride form of zmix_f64 over several problems (element = one (bx, ax)).
__global__ void __launch_bounds__(kThreads) zmix_grouped_kernel(const GroupedElem* __restrict__ probs, int nprob,
                                                               const int* __restrict__ nprob_dev) {
  if (nprob_dev) nprob = *nprob_dev;
  if (nprob <= 0) return;
  const GroupedElem& last = probs[nprob - 1];
  const int64_t total = last.begin + (int64_t)last.Bx * last.Ax;
  int q = 0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kThreads) {
    q = find_elem(probs, nprob, i, q);
    const GroupedElem& g = probs[q];
    const int64_t li = i - g.begin;
    const int bx = (int)(li % g.Bx);
    const int64_t ax = li / g.Bx;
    const double2* src = g.in + bx + (int64_t)g.Bx * g.nin * ax;
    double2* dst = g.out + bx + (int64_t)g.Bx * g.nout * ax;
    for (int o0 = 0; o0 < g.nout; o0 += kMixMax) {
      double2 acc[kMixMax];
#pragma unroll
      for (int o = 0; o < kMixMax; ++o) acc[o] = double2{0.0, 0.0};
      for (int k = 0; k < g.nin; ++k) {
        const double2 v = src[(int64_t)g.Bx * k];
#pragma unroll
        for (int o = 0; o < kMixMax; ++o) {
          if (o0 + o < g.nout) {
            const double2 w = g.aux[o0 + o + g.nout * k];
            acc[o].x = fma(w.x, v.x, fma(-w.y, v.y, acc[o].x));
            acc[o].y = fma(w.x, v.y, fma(w.y, v.x, acc[o].y));
          }
        }
      }
#pragma unroll
      for (int o = 0; o < kMixMax; ++o)
        if (o0 + o < g.nout) dst[(int64_t)g.Bx * (o0 + o)] = acc[o];
    }
  }
}

__global__ void __launch_bounds__(kThreads) zsum_grouped_kernel(const GroupedElem* __restrict__ probs, int nprob,
                                                               const int* __restrict__ nprob_dev) {
  if (nprob_dev) nprob = *nprob_dev;
  if (nprob <= 0) return;
  const GroupedElem& last = probs[nprob - 1];
  const int64_t total = last.begin + last.n;
  int q = 0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kThreads) {
    q = find_elem(probs, nprob, i, q);
    const GroupedElem& g = probs[q];
    const int64_t li = i - g.begin;
    double2 acc = g.in[li];
    for (int k = 1; k < g.np; ++k) {
      const double2 v = g.in[(int64_t)k * g.stride + li];
      acc.x += v.x;
      acc.y += v.y;
    }
    if (g.aux) {
      const double2 v = g.aux[li];
      acc.x = fma(g.beta, v.x, acc.x);
      acc.y = fma(g.beta, v.y, acc.y);
    }
    g.out[li] = acc;
  }
}
}  // namespace

hipError_t launch_zmix_grouped(const GroupedElem* d, int n, int grid, hipStream_t s, const int* nprob_dev) {
  if ((n <= 0 && !nprob_dev) || grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(zmix_grouped_kernel, dim3(grid), dim3(kThreads), 0, s, d, n, nprob_dev);
  return hipGetLastError();
}

hipError_t launch_zsum_grouped(const GroupedElem* d, int n, int grid, hipStream_t s, const int* nprob_dev) {
  if ((n <= 0 && !nprob_dev) || grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(zsum_grouped_kernel, dim3(grid), dim3(kThreads), 0, s, d, n, nprob_dev);
  return hipGetLastError();
}

hipError_t launch_zmix_f64(const double2* in, double2* out, const double2* mix, int Bx, int nin,
                           int nout, int64_t Ax, hipStream_t s) {
  const int64_t total = (int64_t)Bx * Ax;
  if (total <= 0) return hipSuccess;
  const int64_t blocks = (total + kThreads - 1) / kThreads;
  for (int o0 = 0; o0 < nout; o0 += kMixMax) {
    const int no = min(kMixMax, nout - o0);
    hipLaunchKernelGGL(zmix_f64_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, in,
                       out + (int64_t)Bx * o0, mix + o0, Bx, nin, no, nout, nout, total);
  }
  return hipGetLastError();
}

namespace {

constexpr int kChainT = 1024;

// One BRGEMM of a chain: 16 x 16 output blocks (per batch) over the waves, each block the
// sequence of v_mfma_f64_16x16x4f64 the tiled kernel issues for it (batches at s*b, r outer,
// k ascending in steps of 4 with lane group q on k0 + q, zero beyond M, N, K; the conjugate
// and the three-product operands formed as its LDS staging forms them) and the same epilogue.
__device__ void chain_gemm(const ZgemmDesc& p) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lc = lane & 15, q = lane >> 4;
  const int gm = (p.M + 15) / 16, gn = (p.N + 15) / 16, per_b = gm * gn;
  const double csign = p.conjA ? -1.0 : 1.0;
  for (int blk = wave; blk < per_b * p.NB; blk += kChainT / 64) {
    const int b = blk / per_b, rem = blk - b * per_b;
    const int m = (rem % gm) * 16 + lc, n = (rem / gm) * 16 + lc;
    const double2* Ab = p.A + (int64_t)b * p.sAb;
    const double2* Bb = p.B + (int64_t)b * p.sBb;
    d4 p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
    for (int r = 0; r < p.R; ++r) {
#pragma unroll 4
      for (int k0 = 0; k0 < p.K; k0 += 4) {
        const int k = k0 + q;
        const double2 av = (m < p.M && k < p.K) ? Ab[m * p.sAm + k * p.sAk + r * p.sAr] : double2{0.0, 0.0};
        const double2 bv = (k < p.K && n < p.N) ? Bb[k * p.sBk + n * p.sBn + r * p.sBr] : double2{0.0, 0.0};
        const double ar = av.x, ai = csign * av.y;
        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bv.x, p1, 0, 0, 0);
        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bv.y, p2, 0, 0, 0);
        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, bv.x + bv.y, p3, 0, 0, 0);
      }
    }
    double2* Cb = p.C + (int64_t)b * p.sCb;
    const double2* Db = p.Dadd ? p.Dadd + (int64_t)b * p.sCb : nullptr;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int mo = (rem % gm) * 16 + q + 4 * rr;
      if (mo < p.M && n < p.N) {
        const int64_t off = mo * p.sCm + n * p.sCn;
        double2 v{p1[rr] - p2[rr], p3[rr] - (p1[rr] + p2[rr])};
        if (Db) {
          const double2 x = Db[off];
          v.x = fma(p.beta, x.x, v.x);
          v.y = fma(p.beta, x.y, v.y);
        }
        Cb[off] = v;
      }
    }
  }
}

// One MPO mix of a chain, element by element as zmix_f64_kernel sums it.
__device__ void chain_mix(const ChainStep& c) {
  const int64_t total = (int64_t)c.Bx * c.Ax;
  for (int64_t gid = threadIdx.x; gid < total; gid += kChainT) {
    const int bx = (int)(gid % c.Bx);
    const int64_t ax = gid / c.Bx;
    const double2* src = c.in + bx + (int64_t)c.Bx * c.nin * ax;
    double2* dst = c.out + bx + (int64_t)c.Bx * c.nout * ax;
    for (int o0 = 0; o0 < c.nout; o0 += kMixMax) {
      double2 acc[kMixMax];
#pragma unroll
      for (int o = 0; o < kMixMax; ++o) acc[o] = double2{0.0, 0.0};
      for (int i = 0; i < c.nin; ++i) {
        const double2 v = src[(int64_t)c.Bx * i];
#pragma unroll
        for (int o = 0; o < kMixMax; ++o) {
          if (o0 + o < c.nout) {
            const double2 w = c.mix[o0 + o + c.nout * i];
            acc[o].x = fma(w.x, v.x, fma(-w.y, v.y, acc[o].x));
            acc[o].y = fma(w.x, v.y, fma(w.y, v.x, acc[o].y));
          }
        }
      }
#pragma unroll
      for (int o = 0; o < kMixMax; ++o)
        if (o0 + o < c.nout) dst[(int64_t)c.Bx * (o0 + o)] = acc[o];
    }
  }
}

__global__ void __launch_bounds__(kChainT) zchain_kernel(const Chain c) {
  for (int i = 0; i < c.nsteps; ++i) {
    if (c.step[i].kind == 0)
      chain_gemm(c.step[i].g);
    else
      chain_mix(c.step[i]);
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_zchain_f64(const Chain& c, hipStream_t s) {
  if (c.nsteps < 1 || c.nsteps > 3) return hipErrorInvalidValue;
  for (int i = 0; i < c.nsteps; ++i)
    if (c.step[i].kind == 0 && (c.step[i].g.arith != 0 || c.step[i].g.group != 0)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(zchain_kernel, dim3(1), dim3(kChainT), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_heff_full_f64(const double2* L, const double2* W, const double2* R, double2* H,
                                int d, int Db, int wl, int Dk, int Dbr, int wr, int Dkr,
                                double offset, hipStream_t s) {
  const int64_t nrows = (int64_t)d * Db * Dbr, ncols = (int64_t)d * Dk * Dkr;
  const int64_t total = nrows * ncols;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(heff_full_f64_kernel, dim3((unsigned)((total + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, s, L, W, R, H, d, Db, wl, Dk, Dbr, wr, Dkr, offset, nrows,
                     total);
  return hipGetLastError();
}

}  // namespace mpsd
