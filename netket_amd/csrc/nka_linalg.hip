// Dense linear algebra of the SR / minSR solve on gfx950 matrix cores.
//
// nka_gram_bf16x3  S = O_L^T O_L or T = O_L O_L^T (+ shift I + offset 11^T)
//                  (netket/optimizer/qgt/qgt_jacobian_dense.py:136-161,
//                   netket/_src/ngd/sr.py:58-63, netket/_src/ngd/srt.py:55-66)
//                  from a bf16 hi/lo split with three v_mfma_f32_32x32x16_bf16
//                  products per tile (hi.hi + hi.lo + lo.hi), FP32 accumulate.
// nka_gemm_nt_f64 / nka_syrk_f64
//                  FP64 v_mfma_f64_16x16x4 products, used by the Cholesky
//                  trailing update, its panel solve and FP64 Gram matrices.
// nka_potrf_f64 / nka_potrs_f64
//                  blocked right-looking Cholesky and the two triangular
//                  solves (jax.scipy.linalg.cho_factor / cho_solve as called by
//                  netket/optimizer/solver/solvers.py:241-270).
// nka_gemv_f64, nka_check_finite, nka_sgd_update, nka_centre_scale, nka_stats_moments
//                  vector helpers of the step.
#include "nka_common.hpp"
#include "netket_amd.h"
#include <mutex>

namespace nka {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// Map a linear block id onto the lower-triangular tile pair (ti >= tj).
__device__ inline void tri_tile(int64_t b, int& ti, int& tj) {
  int64_t i = (int64_t)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= b) ++i;
  while (i * (i + 1) / 2 > b) --i;
  ti = (int)i;
  tj = (int)(b - i * (i + 1) / 2);
}

// Spread consecutive block ids over the 8 XCDs in contiguous chunks: the
// hardware dispatches block b to XCD b % 8, so block b is given the work item
// (b % 8) * ceil(G/8) + b / 8; neighbouring tiles (which share operand rows)
// then run on the same XCD and share its L2.
// The grid is launched with a multiple of 8 blocks (G = 8 * per), which makes
// the map a bijection of [0, G); ids past the real work count exit early.
__device__ inline int64_t xcd_remap(int64_t b, int64_t G) {
  const int64_t per = G / 8;
  return (b % 8) * per + b / 8;
}
inline unsigned grid_mult8(int64_t tiles) { return (unsigned)(((tiles + 7) / 8) * 8); }

// ---------------------------------------------------------------------------
// bf16x3 Gram: 128x128 output tile per workgroup, 4 waves of 64x64, K step 64.
// ---------------------------------------------------------------------------
constexpr int GB_K = 64;    // K step (bf16 elements): one 128-byte line per row and plane
constexpr int GB_LD = 72;   // LDS row stride in bf16 (64 + 8 pad: 144 bytes)
constexpr int GB_CPR = GB_K / 8;              // 16-byte chunks per row and plane

// Epilogue of the RBM neural tangent kernel, T = O_c O_c^T scaled, from
// Theta = t t^T (the bf16x3 product) and S = sigma sigma^T:
//   T_ij = scale ((hb + S_ij) Theta_ij + ha S_ij - v_i - v_j + c) + offset
// (+ shift on the diagonal), with S_ij = N - 2 popcount(bits_i ^ bits_j)
// exact from the packed spins.
struct GramRbmEpi {
  const uint32_t* bits;   // (n, nw): bit b of word w is (sigma_{32w+b} > 0)
  int nw, nwp, N;         // words per row, LDS row stride (multiple of 4), sites
  const double* v;
  const double* c;
  double hb, ha, scale;
  // two-layer network (NP = 3): the passes 1 and 2 take these planes, and
  //   T_ij = scale ((hb2 + H_ij) G_ij + (hb + X_ij) D_ij - v_i - v_j + c) + offset
  // with H = h1 h1^T (pass 0), G = g2 g2^T (pass 1), D = d1 d1^T (pass 2),
  // X = x x^T from the packed spins
  const uint16_t* xh1;
  const uint16_t* xl1;
  const uint16_t* xh2;
  const uint16_t* xl2;
  int64_t K1, K2, ld1, ld2;
  double hb2;
};

// WT: wave tile edge (64 or 96); the workgroup tile is 2 WT x 2 WT (four
// waves), so each wave holds (WT/32)^2 accumulators of 32x32.  A 192-wide
// tile reads 12 operand fragments from LDS per 27 products (128-wide: 8 per
// 12) and re-reads each operand row panel from L2 a third fewer times; it
// needs 110 KB of LDS, so one workgroup runs per CU.
// NP: products per tile, 1 (the Gram, or the RBM tangent kernel when RBM) or
// 3 (the two-layer network's tangent kernel, RBM set), each K loop in turn.
template <bool RBM, int WT, int NP = 1>
__global__ void __launch_bounds__(256, 1) gram_bf16x3_kernel(const uint16_t* __restrict__ xh,
                                                          const uint16_t* __restrict__ xl, int64_t n,
                                                          int64_t K, int64_t ldx, double shift,
                                                          double offset, double* __restrict__ out,
                                                          int64_t ldo, int64_t n_tiles_total,
                                                          GramRbmEpi ep) {
  constexpr int TT = 2 * WT;                 // workgroup tile edge
  constexpr int MI = WT / 32;                // 32x32 tiles per wave edge
  constexpr int NIT = TT * GB_CPR / 256;     // 16-byte chunks per thread and plane
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][TT * GB_LD];  // [plane][row*LD + k]
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][TT * GB_LD];
  const int64_t bid = xcd_remap(blockIdx.x, gridDim.x);
  if (bid >= n_tiles_total) return;
  int ti, tj;
  tri_supertile(bid, (n + TT - 1) / TT, ti, tj);
  const int64_t i0 = (int64_t)ti * TT, j0 = (int64_t)tj * TT;
  const int tid = threadIdx.x, wid = tid / 64, lane = lane_id();
  const int wm = (wid >> 1) * WT, wn = (wid & 1) * WT;  // wave tile origin in block tile
  const bool diag_tile = (ti == tj);

  floatx16 accs[NP][MI][MI];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MI; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) accs[p][a][b][e] = 0.0f;

#pragma unroll
  for (int pass = 0; pass < NP; ++pass) {
    const uint16_t* xh_p = pass == 0 ? xh : pass == 1 ? ep.xh1 : ep.xh2;
    const uint16_t* xl_p = pass == 0 ? xl : pass == 1 ? ep.xl1 : ep.xl2;
    const int64_t K_p = pass == 0 ? K : pass == 1 ? ep.K1 : ep.K2;
    const int64_t ld_p = pass == 0 ? ldx : pass == 1 ? ep.ld1 : ep.ld2;
    if (pass > 0) __syncthreads();  // the previous pass's last slab has been read
    // Each operand slab is TT rows x GB_K k = TT GB_CPR chunks of 8 bf16;
    // thread tid moves chunks tid + 256 it (it < NIT) of each of the (up to)
    // four planes.  The next slab is fetched into registers while the current
    // one is used.
    uint4 pf[NIT][4];
    // chunks it0 <= it < it1 of the slab at k0
    auto fetch_part = [&](int64_t k0, int it0, int it1) {
  #pragma unroll
      for (int it = 0; it < NIT; ++it) {
        if (it < it0 || it >= it1) continue;
        const int chunk = tid + it * 256;
        const int r = chunk / GB_CPR, kc = (chunk % GB_CPR) * 8;
        const int64_t kk = k0 + kc;
        uint4 z = {0, 0, 0, 0};
        pf[it][0] = z;
        pf[it][1] = z;
        pf[it][2] = z;
        pf[it][3] = z;
        if (i0 + r < n && kk < K_p) {
          pf[it][0] = *(const uint4*)(xh_p + (i0 + r) * ld_p + kk);
          pf[it][1] = *(const uint4*)(xl_p + (i0 + r) * ld_p + kk);
        }
        if (!diag_tile && j0 + r < n && kk < K_p) {
          pf[it][2] = *(const uint4*)(xh_p + (j0 + r) * ld_p + kk);
          pf[it][3] = *(const uint4*)(xl_p + (j0 + r) * ld_p + kk);
        }
      }
    };
    fetch_part(0, 0, NIT);
    for (int64_t k0 = 0; k0 < K_p; k0 += GB_K) {
      __syncthreads();
  #pragma unroll
      for (int it = 0; it < NIT; ++it) {
        const int chunk = tid + it * 256;
        const int r = chunk / GB_CPR, kc = (chunk % GB_CPR) * 8;
        *(uint4*)(&sA[0][r * GB_LD + kc]) = pf[it][0];
        *(uint4*)(&sA[1][r * GB_LD + kc]) = pf[it][1];
        if (!diag_tile) {
          *(uint4*)(&sB[0][r * GB_LD + kc]) = pf[it][2];
          *(uint4*)(&sB[1][r * GB_LD + kc]) = pf[it][3];
        }
      }
      __syncthreads();
      const uint16_t* Bh = diag_tile ? sA[0] : sB[0];
      const uint16_t* Bl = diag_tile ? sA[1] : sB[1];
      const bool more = k0 + GB_K < K_p;
  #pragma unroll
      for (int ks = 0; ks < GB_K; ks += 16) {
        // the next slab's loads are spread over the first K sub-steps, so the
        // memory pipeline works through them while the products run
        if (more) {
          constexpr int per = (NIT + GB_K / 16 - 2) / (GB_K / 16 - 1);
          const int q = ks / 16;
          if (q < GB_K / 16 - 1) fetch_part(k0 + GB_K, q * per, (q + 1) * per < NIT ? (q + 1) * per : NIT);
        }
        __builtin_amdgcn_sched_barrier(0);
        const int kq = ks + (lane >> 5) * 8;
        bf16x8 ah[MI], al[MI], bh[MI], bl[MI];
  #pragma unroll
        for (int a = 0; a < MI; ++a) {
          const int row = wm + a * 32 + (lane & 31);
          ah[a] = *(const bf16x8*)(&sA[0][row * GB_LD + kq]);
          al[a] = *(const bf16x8*)(&sA[1][row * GB_LD + kq]);
        }
  #pragma unroll
        for (int b = 0; b < MI; ++b) {
          const int row = wn + b * 32 + (lane & 31);
          bh[b] = *(const bf16x8*)(&Bh[row * GB_LD + kq]);
          bl[b] = *(const bf16x8*)(&Bl[row * GB_LD + kq]);
        }
        // small products first, four independent accumulators between
        // consecutive updates of the same accumulator
  #pragma unroll
        for (int a = 0; a < MI; ++a)
  #pragma unroll
          for (int b = 0; b < MI; ++b)
            accs[pass][a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], accs[pass][a][b], 0, 0, 0);
  #pragma unroll
        for (int a = 0; a < MI; ++a)
  #pragma unroll
          for (int b = 0; b < MI; ++b)
            accs[pass][a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], accs[pass][a][b], 0, 0, 0);
  #pragma unroll
        for (int a = 0; a < MI; ++a)
  #pragma unroll
          for (int b = 0; b < MI; ++b)
            accs[pass][a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], accs[pass][a][b], 0, 0, 0);
      }
    }
  }
  // store: element e of lane l in a 32x32 tile is (row (e/4)*8 + (l/32)*4 + e%4, col l%32)
  if constexpr (!RBM) {
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MI; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int64_t gi = i0 + wm + a * 32 + (e / 4) * 8 + (lane >> 5) * 4 + (e & 3);
          const int64_t gj = j0 + wn + b * 32 + (lane & 31);
          if (gi < n && gj <= gi) {
            double v = (double)accs[0][a][b][e] + offset;
            if (gi == gj) v += shift;
            out[gi * ldo + gj] = v;
          }
        }
  } else {
    // spins of the tile's 128 rows and 128 columns into the operand LDS
    __syncthreads();
    uint32_t* rb = (uint32_t*)&sA[0][0];
    uint32_t* cb = (uint32_t*)&sB[0][0];
    const int nwp = ep.nwp;
    for (int idx = tid; idx < TT * nwp; idx += 256) {
      const int r = idx / nwp, w = idx % nwp;
      uint32_t x = 0, y = 0;
      if (w < ep.nw) {
        if (i0 + r < n) x = ep.bits[(i0 + r) * ep.nw + w];
        if (j0 + r < n) y = ep.bits[(j0 + r) * ep.nw + w];
      }
      rb[idx] = x;
      cb[idx] = y;
    }
    __syncthreads();
    int cnt[MI][MI][16];
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MI; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) cnt[a][b][e] = 0;
    for (int w = 0; w < nwp; w += 4) {
      uint4 cq[MI];
#pragma unroll
      for (int b = 0; b < MI; ++b) cq[b] = *(const uint4*)&cb[(wn + b * 32 + (lane & 31)) * nwp + w];
#pragma unroll
      for (int a = 0; a < MI; ++a)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int rl = wm + a * 32 + (e / 4) * 8 + (lane >> 5) * 4 + (e & 3);
          const uint4 rq = *(const uint4*)&rb[rl * nwp + w];
#pragma unroll
          for (int b = 0; b < MI; ++b)
            cnt[a][b][e] += __builtin_popcount(rq.x ^ cq[b].x) + __builtin_popcount(rq.y ^ cq[b].y) +
                            __builtin_popcount(rq.z ^ cq[b].z) + __builtin_popcount(rq.w ^ cq[b].w);
        }
    }
    const double cc = *ep.c;
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < MI; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int64_t gi = i0 + wm + a * 32 + (e / 4) * 8 + (lane >> 5) * 4 + (e & 3);
          const int64_t gj = j0 + wn + b * 32 + (lane & 31);
          if (gi < n && gj <= gi) {
            const double S = (double)(ep.N - 2 * cnt[a][b][e]);
            double v;
            if constexpr (NP == 1) {
              const double Th = (double)accs[0][a][b][e];
              v = ep.scale * ((ep.hb + S) * Th + ep.ha * S - ep.v[gi] - ep.v[gj] + cc) + offset;
            } else {
              const double Hh = (double)accs[0][a][b][e];
              const double Gg = (double)accs[NP - 2][a][b][e];
              const double Dd = (double)accs[NP - 1][a][b][e];
              v = ep.scale * ((ep.hb2 + Hh) * Gg + (ep.hb + S) * Dd - ep.v[gi] - ep.v[gj] + cc) + offset;
            }
            if (gi == gj) v += shift;
            out[gi * ldo + gj] = v;
          }
        }
  }
}

// Largest LDS row stride (words) of the packed spins in the RBM epilogue: the
// tile's rows must fit in one operand buffer (2 planes x rows x GB_LD bf16).
constexpr int kGramRbmMaxNwp = (int)(2 * GB_LD * sizeof(uint16_t) / sizeof(uint32_t)) / 4 * 4;
// wave tile edges of the Gram and of the RBM tangent kernel
constexpr int kGramWT = 96;
constexpr int kGramRbmWT = 64;

// ---------------------------------------------------------------------------
// C = A (Bh + Bl)^T in FP32 (ma x nb), K contiguous in both operands (row
// stride ldk, zero beyond the true K).  A is one bf16 plane whose entries
// bf16 holds exactly (products of +-1 spins, 0 and 1); B is split into two
// bf16 planes.  Each fragment pair therefore takes two products instead of
// the Gram's three, and A's LDS slab is a single plane.  Four waves of
// 32 MIA x 32 MIB (v_mfma_f32_32x32x16_bf16) make a 64 MIA x 64 MIB tile;
// K step, slab prefetch and fragment layout are the Gram kernel's.  Tiles are
// numbered down the columns (the ma side fastest), so the workgroups that
// share a B column panel (the larger operand) sit on one XCD and read it
// through one L2.
// The FP32 accumulation inside the MFMA rounds with a bias (the error of a
// plain sum over 16384 samples grows like the number of accumulations times
// 2^-24, not like its square root), so the products are accumulated on the
// matrix cores for kExactAChunk K steps (128 accumulations) at a time and each
// chunk is added into a separate FP32 total with round-to-nearest vector adds.
// 192 x 128 tiles (3 x 2 blocks of 32 x 32 per wave) leave the registers for
// that total; the single-plane operand takes the 192-row side, so a K step
// reads 7 fragments per 12 products from LDS.
// ---------------------------------------------------------------------------
constexpr int kExactAMIA = 3, kExactAMIB = 2;
constexpr int kExactAChunk = 16;

template <int MIA, int MIB>
__global__ void __launch_bounds__(256, 1) gemm_exact_a_kernel(const uint16_t* __restrict__ a, int64_t ma,
                                                           const uint16_t* __restrict__ bh,
                                                           const uint16_t* __restrict__ bl, int64_t nb,
                                                           int64_t K, int64_t ldk, float* __restrict__ c,
                                                           int64_t ldc, int64_t tiles_m, int64_t n_tiles_total) {
  constexpr int TA = 64 * MIA, TB = 64 * MIB;
  constexpr int NITA = TA * GB_CPR / 256, NITB = TB * GB_CPR / 256;
  constexpr int NIT = NITA > NITB ? NITA : NITB;
  __shared__ __attribute__((aligned(16))) uint16_t sA[TA * GB_LD];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][TB * GB_LD];
  const int64_t bid = xcd_remap(blockIdx.x, gridDim.x);
  if (bid >= n_tiles_total) return;
  const int64_t i0 = (bid % tiles_m) * TA, j0 = (bid / tiles_m) * TB;
  const int tid = threadIdx.x, wid = tid / 64, lane = lane_id();
  const int wm = (wid >> 1) * 32 * MIA, wn = (wid & 1) * 32 * MIB;

  floatx16 acc[MIA][MIB];
  float tot[MIA][MIB][16];
#pragma unroll
  for (int x = 0; x < MIA; ++x)
#pragma unroll
    for (int y = 0; y < MIB; ++y)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        acc[x][y][e] = 0.0f;
        tot[x][y][e] = 0.0f;
      }

  // chunks it0 <= it < it1 of the slab at k0: A, B hi, B lo
  uint4 pa[NITA], pb[NITB][2];
  auto fetch_part = [&](int64_t k0, int it0, int it1) {
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      if (it < it0 || it >= it1) continue;
      const int chunk = tid + it * 256;
      const int r = chunk / GB_CPR, kc = (chunk % GB_CPR) * 8;
      const int64_t kk = k0 + kc;
      const uint4 z = {0, 0, 0, 0};
      if (it < NITA) {
        pa[it] = z;
        if (i0 + r < ma && kk < K) pa[it] = *(const uint4*)(a + (i0 + r) * ldk + kk);
      }
      if (it < NITB) {
        pb[it][0] = z;
        pb[it][1] = z;
        if (j0 + r < nb && kk < K) {
          pb[it][0] = *(const uint4*)(bh + (j0 + r) * ldk + kk);
          pb[it][1] = *(const uint4*)(bl + (j0 + r) * ldk + kk);
        }
      }
    }
  };
  fetch_part(0, 0, NIT);
  for (int64_t k0 = 0; k0 < K; k0 += GB_K) {
    if (k0 != 0 && k0 % (GB_K * kExactAChunk) == 0) {
#pragma unroll
      for (int x = 0; x < MIA; ++x)
#pragma unroll
        for (int y = 0; y < MIB; ++y)
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            tot[x][y][e] += acc[x][y][e];
            acc[x][y][e] = 0.0f;
          }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int chunk = tid + it * 256;
      const int r = chunk / GB_CPR, kc = (chunk % GB_CPR) * 8;
      if (it < NITA) *(uint4*)(&sA[r * GB_LD + kc]) = pa[it];
      if (it < NITB) {
        *(uint4*)(&sB[0][r * GB_LD + kc]) = pb[it][0];
        *(uint4*)(&sB[1][r * GB_LD + kc]) = pb[it][1];
      }
    }
    __syncthreads();
    const bool more = k0 + GB_K < K;
#pragma unroll
    for (int ks = 0; ks < GB_K; ks += 16) {
      // the next slab's loads spread over the first K sub-steps
      if (more) {
        constexpr int per = (NIT + GB_K / 16 - 2) / (GB_K / 16 - 1);
        const int q = ks / 16;
        if (q < GB_K / 16 - 1) fetch_part(k0 + GB_K, q * per, (q + 1) * per < NIT ? (q + 1) * per : NIT);
      }
      __builtin_amdgcn_sched_barrier(0);
      const int kq = ks + (lane >> 5) * 8;
      bf16x8 af[MIA], bhf[MIB], blf[MIB];
#pragma unroll
      for (int x = 0; x < MIA; ++x) af[x] = *(const bf16x8*)(&sA[(wm + x * 32 + (lane & 31)) * GB_LD + kq]);
#pragma unroll
      for (int y = 0; y < MIB; ++y) {
        const int row = wn + y * 32 + (lane & 31);
        bhf[y] = *(const bf16x8*)(&sB[0][row * GB_LD + kq]);
        blf[y] = *(const bf16x8*)(&sB[1][row * GB_LD + kq]);
      }
      // small products first
#pragma unroll
      for (int x = 0; x < MIA; ++x)
#pragma unroll
        for (int y = 0; y < MIB; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[x], blf[y], acc[x][y], 0, 0, 0);
#pragma unroll
      for (int x = 0; x < MIA; ++x)
#pragma unroll
        for (int y = 0; y < MIB; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[x], bhf[y], acc[x][y], 0, 0, 0);
    }
  }
  // element e of lane l in a 32x32 tile is (row (e/4)*8 + (l/32)*4 + e%4, col l%32)
#pragma unroll
  for (int x = 0; x < MIA; ++x)
#pragma unroll
    for (int y = 0; y < MIB; ++y)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t gi = i0 + wm + x * 32 + (e / 4) * 8 + (lane >> 5) * 4 + (e & 3);
        const int64_t gj = j0 + wn + y * 32 + (lane & 31);
        if (gi < ma && gj < nb) c[gi * ldc + gj] = tot[x][y][e] + acc[x][y][e];
      }
}

// ---------------------------------------------------------------------------
// FP64 NT GEMM: C = beta C + alpha A B^T (+ shift on the diagonal),
// 64x64 tile per workgroup, 4 waves of 32x32, K step 16.  lower_only skips
// tiles above the diagonal and elements with j > i.  A and C may alias when
// beta == 0 and the grid has a single column tile (in-place panel solve):
// every workgroup reads all of its A rows before it writes them.
// ---------------------------------------------------------------------------
constexpr int GD_T = 64;
constexpr int GD_K = 16;
constexpr int GD_LD = 17;  // padded LDS stride (doubles)

__global__ void __launch_bounds__(256) gemm_nt_f64_kernel(const double* a, int64_t lda, const double* b,
                                                          int64_t ldb, double* c, int64_t ldc, int64_t m,
                                                          int64_t n, int64_t K, double alpha, double beta,
                                                          double shift, int lower_only, int64_t tiles_n,
                                                          int64_t n_tiles_total, const int32_t* info) {
  if (info && *info != 0) return;
  __shared__ double sA[GD_T * GD_LD];
  __shared__ double sB[GD_T * GD_LD];
  const int64_t bid = xcd_remap(blockIdx.x, gridDim.x);
  if (bid >= n_tiles_total) return;
  int ti, tj;
  if (lower_only) {
    tri_tile(bid, ti, tj);
  } else {
    ti = (int)(bid / tiles_n);
    tj = (int)(bid % tiles_n);
  }
  const int64_t i0 = (int64_t)ti * GD_T, j0 = (int64_t)tj * GD_T;
  const int tid = threadIdx.x, wid = tid / 64, lane = lane_id();
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  doublex4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = (doublex4){0.0, 0.0, 0.0, 0.0};

  for (int64_t k0 = 0; k0 < K; k0 += GD_K) {
    __syncthreads();
    // 64 rows x 16 k per operand = 1024 doubles; 256 threads x 4
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + it * 256;
      const int r = idx >> 4, kc = idx & 15;
      const int64_t kk = k0 + kc;
      sA[r * GD_LD + kc] = (i0 + r < m && kk < K) ? a[(i0 + r) * lda + kk] : 0.0;
      sB[r * GD_LD + kc] = (j0 + r < n && kk < K) ? b[(j0 + r) * ldb + kk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < GD_K; ks += 4) {
      const int kq = ks + (lane >> 4);
      double av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = sA[(wm + x * 16 + (lane & 15)) * GD_LD + kq];
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = sB[(wn + y * 16 + (lane & 15)) * GD_LD + kq];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
  }
  __syncthreads();
  // element e of lane l in a 16x16 tile is (row l/16 + 4e, col l%16); C is
  // read before any of it is written, so the loads' latencies overlap
  double cv[2][2][4];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t gi = i0 + wm + x * 16 + (lane >> 4) + 4 * e;
        const int64_t gj = j0 + wn + y * 16 + (lane & 15);
        cv[x][y][e] = (beta != 0.0 && gi < m && gj < n && (!lower_only || gj <= gi)) ? c[gi * ldc + gj] : 0.0;
      }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t gi = i0 + wm + x * 16 + (lane >> 4) + 4 * e;
        const int64_t gj = j0 + wn + y * 16 + (lane & 15);
        if (gi < m && gj < n && (!lower_only || gj <= gi)) {
          double v = alpha * acc[x][y][e];
          if (beta != 0.0) v += beta * cv[x][y][e];
          if (gi == gj) v += shift;
          c[gi * ldc + gj] = v;
        }
      }
}

// ---------------------------------------------------------------------------
// Large FP64 NT GEMM for the trailing updates: C = beta C + alpha A B^T on a
// 128x128 tile per workgroup, 4 waves of 64x64 (4x4 v_mfma_f64_16x16x4
// blocks, 16 accumulators per wave), K step 16.  The next K tile is loaded
// into registers while the current one is multiplied out of LDS; the LDS
// tiles are k-major so a fragment read is 16 consecutive rows.  lower_only
// skips tiles above the diagonal and elements with j > i.  A, B and C must
// not overlap.
// ---------------------------------------------------------------------------
constexpr int G2_T = 128;
constexpr int G2_K = 16;
constexpr int G2_LD = G2_T + 4;   // doubles per k row of an LDS tile

template <bool VEC>
__device__ __forceinline__ void gemm_big_tile(const double* __restrict__ a, int64_t lda, const double* __restrict__ b,
                                     int64_t ldb, double* __restrict__ c, int64_t ldc, int64_t m, int64_t n,
                                     int64_t K, double alpha, double beta, int lower_only, int64_t tiles_n,
                                     int64_t bid, double* sA, double* sB, double* __restrict__ wout) {
  int ti, tj;
  if (lower_only) {
    tri_supertile(bid, tiles_n, ti, tj);
  } else {
    ti = (int)(bid / tiles_n);
    tj = (int)(bid % tiles_n);
  }
  const int64_t i0 = (int64_t)ti * G2_T, j0 = (int64_t)tj * G2_T;
  const int tid = threadIdx.x, wid = tid / 64, lane = lane_id();
  const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
  // global -> register mapping: thread t loads row t/2, k = (t%2)*8 .. +8 of
  // each operand.  Rows past the edge re-read the last row (K is a multiple
  // of 16, so every load is in bounds); their products are never stored.
  const int lr = tid >> 1, lk = (tid & 1) * 8;
  const double* pa = a + min<int64_t>(i0 + lr, m - 1) * lda + lk;
  const double* pb = b + min<int64_t>(j0 + lr, n - 1) * ldb + lk;
  double ra[8], rb[8];
  auto load = [&](int64_t k0) {
    if (VEC) {
#pragma unroll
      for (int q = 0; q < 8; q += 2) {
        const double2 va = *reinterpret_cast<const double2*>(pa + k0 + q);
        const double2 vb = *reinterpret_cast<const double2*>(pb + k0 + q);
        ra[q] = va.x; ra[q + 1] = va.y;
        rb[q] = vb.x; rb[q + 1] = vb.y;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        ra[q] = pa[k0 + q];
        rb[q] = pb[k0 + q];
      }
    }
  };
  doublex4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = (doublex4){0.0, 0.0, 0.0, 0.0};
  load(0);
  for (int64_t k0 = 0; k0 < K; k0 += G2_K) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      sA[(lk + q) * G2_LD + lr] = ra[q];
      sB[(lk + q) * G2_LD + lr] = rb[q];
    }
    __syncthreads();
    if (k0 + G2_K < K) load(k0 + G2_K);
#pragma unroll
    for (int ks = 0; ks < G2_K; ks += 4) {
      const int kq = ks + (lane >> 4);
      double av[4], bv[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) av[x] = sA[kq * G2_LD + wm + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 4; ++y) bv[y] = sB[kq * G2_LD + wn + y * 16 + (lane & 15)];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
  }
  // element e of lane l in a 16x16 block is (row l/16 + 4e, col l%16)
  if (wout) {
    // partial product of a K slice: the raw 128x128 tile, row-major
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          wout[(wm + x * 16 + (lane >> 4) + 4 * e) * G2_T + wn + y * 16 + (lane & 15)] = acc[x][y][e];
    return;
  }
  // C is read in batches of one 16-row block (16 values per lane) before
  // any of the batch is written, and the batch's results are all formed
  // before its stores: the compiler cannot prove the addresses distinct for a
  // runtime ldc, so element-wise read-modify-write waited for each store.
  // Interior tiles (inside C and, for lower_only, below the diagonal) take
  // unpredicated loads and stores, which keep each store's data in its own
  // registers; edge tiles check every element.
  const bool interior = i0 + G2_T <= m && j0 + G2_T <= n && (!lower_only || j0 + G2_T <= i0);
#pragma unroll
  for (int xb = 0; xb < 4; ++xb) {
    double cv[1][4][4];
#pragma unroll
    for (int x = 0; x < 1; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int e = 0; e < 4; ++e) cv[x][y][e] = 0.0;
    if (beta != 0.0) {
#pragma unroll
      for (int x = 0; x < 1; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t gi = i0 + wm + (xb + x) * 16 + (lane >> 4) + 4 * e;
            const int64_t gj = j0 + wn + y * 16 + (lane & 15);
            if (interior || (gi < m && gj < n && (!lower_only || gj <= gi))) cv[x][y][e] = c[gi * ldc + gj];
          }
    }
#pragma unroll
    for (int x = 0; x < 1; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int e = 0; e < 4; ++e) cv[x][y][e] = alpha * acc[xb + x][y][e] + beta * cv[x][y][e];
    if (interior) {
#pragma unroll
      for (int x = 0; x < 1; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            c[(i0 + wm + (xb + x) * 16 + (lane >> 4) + 4 * e) * ldc + j0 + wn + y * 16 + (lane & 15)] = cv[x][y][e];
    } else {
#pragma unroll
      for (int x = 0; x < 1; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t gi = i0 + wm + (xb + x) * 16 + (lane >> 4) + 4 * e;
            const int64_t gj = j0 + wn + y * 16 + (lane & 15);
            if (gi < m && gj < n && (!lower_only || gj <= gi)) c[gi * ldc + gj] = cv[x][y][e];
          }
    }
  }
}


// One tile per workgroup (a grid smaller than the tile count makes each
// workgroup walk tiles bid, bid + G, ..., consecutive in the super-block order).
// SPLIT = false: one tile per unit u < n_units (tile ids from 0), whole K,
// C updated.  SPLIT = true: unit u is slice u % S (kc columns of K) of tile
// n_main + u / S, whose raw product goes to work + u * 128 * 128 for
// gemm_big_reduce_kernel.
template <bool VEC, bool SPLIT>
__global__ void __launch_bounds__(256, 2) gemm_nt_f64_big_kernel(const double* __restrict__ a, int64_t lda,
                                                              const double* __restrict__ b, int64_t ldb,
                                                              double* __restrict__ c, int64_t ldc, int64_t m,
                                                              int64_t n, int64_t K, double alpha, double beta,
                                                              int lower_only, int64_t tiles_n,
                                                              int64_t n_units, int64_t n_main, int S,
                                                              int64_t kc, double* __restrict__ work,
                                                              const int32_t* info) {
  if (info && *info != 0) return;
  __shared__ double sA[G2_K * G2_LD];
  __shared__ double sB[G2_K * G2_LD];
  const int64_t G = gridDim.x;
  for (int64_t u = xcd_remap(blockIdx.x, G); u < n_units; u += G) {
    if (!SPLIT) {
      gemm_big_tile<VEC>(a, lda, b, ldb, c, ldc, m, n, K, alpha, beta, lower_only, tiles_n, u, sA, sB, nullptr);
    } else {
      const int64_t tile = n_main + u / S;
      const int64_t kb = (u % S) * kc, kl = min<int64_t>(kc, K - kb);
      gemm_big_tile<VEC>(a + kb, lda, b + kb, ldb, c, ldc, m, n, kl, alpha, beta, lower_only, tiles_n, tile,
                         sA, sB, work + u * (int64_t)(G2_T * G2_T));
    }
    __syncthreads();
  }
}

// C = beta C + alpha sum_part W[part] for the split tail tiles, parts summed
// in a fixed order.  One workgroup per (tail tile, 32-row band).
__global__ void __launch_bounds__(256) gemm_big_reduce_kernel(const double* __restrict__ work, double* __restrict__ c,
                                                              int64_t ldc, int64_t m, int64_t n, double alpha,
                                                              double beta, int lower_only, int64_t tiles_n,
                                                              int64_t n_main, int S, const int32_t* info) {
  if (info && *info != 0) return;
  const int64_t r = blockIdx.x / 4;  // tail tile index
  const int band = blockIdx.x % 4;
  const int64_t tile = n_main + r;
  int ti, tj;
  if (lower_only) {
    tri_supertile(tile, tiles_n, ti, tj);
  } else {
    ti = (int)(tile / tiles_n);
    tj = (int)(tile % tiles_n);
  }
  const double* w = work + r * S * (int64_t)(G2_T * G2_T);
  // 16 elements per thread, the loads of a slice (and of C) all in flight;
  // each element sums its slices in slice order
  constexpr int PER = 32 * G2_T / 256;
  double acc[PER], cv[PER];
  bool keep[PER];
  int64_t off[PER];
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * 256;
    const int li = band * 32 + idx / G2_T, lj = idx % G2_T;
    const int64_t gi = (int64_t)ti * G2_T + li, gj = (int64_t)tj * G2_T + lj;
    keep[e] = gi < m && gj < n && (!lower_only || gj <= gi);
    off[e] = keep[e] ? gi * ldc + gj : 0;
    acc[e] = 0.0;
    cv[e] = (beta != 0.0 && keep[e]) ? c[off[e]] : 0.0;
  }
  for (int p = 0; p < S; ++p) {
    const double* wp = w + (int64_t)p * G2_T * G2_T + band * 32 * G2_T;
#pragma unroll
    for (int e = 0; e < PER; ++e) acc[e] += wp[threadIdx.x + e * 256];
  }
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    if (keep[e]) {
      double v = alpha * acc[e];
      if (beta != 0.0) v += beta * cv[e];
      c[off[e]] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Cholesky building blocks (block size 64).
// ---------------------------------------------------------------------------
constexpr int CH_NB = 64;

constexpr int CH_LT = CH_NB + 2;  // LDS row stride of the transposed factor (16-byte rows)
constexpr int TS_LD = CH_NB + 1;  // LDS row stride of a 64x64 block (doubles)
constexpr int PF_ROWS = 16;  // rows of the panel per wavefront

// Lane `src`'s value of v in every lane (src a compile-time constant after
// unrolling).
__device__ inline double readlane_f64(double v, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
  return __hiloint2double(hi, lo);
}

// 1/sqrt(a) from the hardware estimate (about 2^-23 relative error) refined by
// two Newton-Raphson steps y <- y + y (1/2 - (a/2) y^2), each of which squares
// the relative error.
__device__ inline double rsqrt_refined(double a) {
  const double h = 0.5 * a;
  double y = __builtin_amdgcn_rsq(a);
  y = fma(y, fma(-(h * y), y, 0.5), y);
  y = fma(y, fma(-(h * y), y, 0.5), y);
  return y;
}

// Bounded wait of a wavefront for an LDS flag of its own workgroup set by
// another wavefront (workgroup-scope acquire; every lane of the setter stores
// the flag, so each lane's earlier LDS accesses are released).  The setter
// always runs to its publication, so the wait ends; past the bound the
// factorisation is reported as failed at this step (*info = k0 + 1) and the
// wavefront goes on.  A factorisation already reported as failed (*info set,
// possibly by another workgroup of the launch, after which a wavefront that
// read it at its start has left) ends the wait too: its result is discarded.
constexpr int64_t kLdsSpinLimit = 1LL << 24;
__device__ inline void lds_wait(const int* flag, int32_t* info, int64_t k0) {
  int64_t spins = 0;
  while (__builtin_amdgcn_readfirstlane(__hip_atomic_load(const_cast<int*>(flag), __ATOMIC_ACQUIRE,
                                                          __HIP_MEMORY_SCOPE_WORKGROUP)) == 0) {
    __builtin_amdgcn_s_sleep(1);
    ++spins;
    // (a normal wait ends within a few hundred polls, before *info is read)
    if (spins > 4096 && (spins & 255) == 0 &&
        __builtin_amdgcn_readfirstlane(__hip_atomic_load(info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
      break;
    if (spins > kLdsSpinLimit) {
      if ((threadIdx.x & 63) == 0) atomicCAS(info, 0, (int32_t)(k0 + 1));
      break;
    }
  }
}

// Publish an LDS flag of the workgroup (every lane stores it).
__device__ inline void lds_publish(int* flag) {
  __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Factor a 64x64 block held in LDS (Ab, row-major, zeros above the diagonal)
// in place with one wavefront, lane i = row i, by 16-column sub-panels:
//  * sub-panel p (columns 16p..16p+15, rows 16p..63) is factored
//    right-looking with the 16 values of each row in registers; at step j the
//    pivot a_jj is read from its lane with v_readlane and the column entries
//    a_cj of the sub-panel's rows c are broadcast through LDS (colbuf), and
//    the update is a_ic -= L_ij L_cj = a_ij a_cj / a_jj, with 1/sqrt(a_jj)
//    from the hardware estimate and two Newton steps;
//  * the block's remaining lower part (rows and columns >= 16(p+1)) then
//    receives the rank-16 update on FP64 MFMA (v_mfma_f64_16x16x4).
// On return Ab holds L (exact zeros above the diagonal) and yv[j] = 1/L_jj.
// With `done` (LDS, zeroed beforehand), done[p] is set with a workgroup-scope
// release once columns 16p..16p+15 of L and their yv entries are final.
// With `ready`, the blocks right of the first sub-panel are written by other
// wavefronts during its factorisation: ready[0..2] are awaited before they
// are first touched (the rank-16 update after sub-panel 0).
// A non-positive or non-finite pivot is replaced by 1 so the factor
// completes; the return value is then the first failing column + 1 + k0
// (0 when every pivot is positive).
__device__ __forceinline__ int factor_block_wave(double* Ab, double* colbuf, double* yv, int lane,
                                                 int64_t k0, int* done = nullptr,
                                                 const int* ready = nullptr, int32_t* info = nullptr) {
  int bad = 0;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    double x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = Ab[lane * TS_LD + 16 * p + q];
    double ajj_next = readlane_f64(x[0], 16 * p);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int jj = 16 * p + j;
      // the pivot comes from v_readlane values (the critical path), the
      // rest of the column through LDS while the reciprocal square root is
      // formed
      colbuf[lane] = x[j];
      const double ajj = ajj_next;
      double a1 = 0.0, b1 = 0.0;
      if (j < 15) {
        a1 = readlane_f64(x[j], jj + 1);
        b1 = readlane_f64(x[j + 1], jj + 1);
      }
      wave_lds_sync();
      double cj[16];
#pragma unroll
      for (int q = j + 1; q < 16; ++q) cj[q] = colbuf[16 * p + q];
      const bool ok = (ajj > 0.0) && isfinite(ajj);
      if (!ok && bad == 0) bad = (int)(k0 + jj + 1);
      const double piv = ok ? ajj : 1.0;
      const double y = rsqrt_refined(piv);
      // the next pivot: lane jj + 1 forms x[j+1] - (x[j] y^2) x[j] below with
      // these same operations (one rounding for the multiply-subtract), so
      // the value is the one that lane holds after this step
      const double yy = y * y;
      if (j < 15) ajj_next = fma(-(a1 * yy), a1, b1);
      // every lane scales and updates: in rows jj and above the entries
      // changed lie above the diagonal, which no later pivot or column
      // broadcast reads and the write-back below replaces by zeros; row jj
      // holds a_jj itself (the value the pivot was formed from)
      const double sij = x[j] * yy;
      x[j] *= y;
      if (!ok && lane == jj) x[j] = piv * y;
      if (lane == jj) yv[jj] = y;
#pragma unroll
      for (int q = j + 1; q < 16; ++q) x[q] -= sij * cj[q];
      wave_lds_sync();
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) Ab[lane * TS_LD + 16 * p + q] = (16 * p + q <= lane) ? x[q] : 0.0;
    wave_lds_sync();
    // columns 16p..16p+15 of L and yv[16p..16p+15] are final from here on
    // (the update below writes only later columns)
    if (done != nullptr) lds_publish(&done[p]);
    if (p == 0 && ready != nullptr) {
#pragma unroll
      for (int i = 0; i < 3; ++i) lds_wait(&ready[i], info, k0);
    }
    // rank-16 update of the lower blocks (bi, bj), p < bj <= bi < 4
#pragma unroll
    for (int bi = p + 1; bi < 4; ++bi)
#pragma unroll
      for (int bj = p + 1; bj <= bi; ++bj) {
        doublex4 acc = (doublex4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kb = 0; kb < 16; kb += 4) {
          const int kk = 16 * p + kb + (lane >> 4);
          const double av = Ab[(bi * 16 + (lane & 15)) * TS_LD + kk];
          const double bv = Ab[(bj * 16 + (lane & 15)) * TS_LD + kk];
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
        }
        // element e of lane l is (row l/16 + 4e, col l%16) of the 16x16 block
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = bi * 16 + (lane >> 4) + 4 * e, c = bj * 16 + (lane & 15);
          Ab[r * TS_LD + c] -= acc[e];
        }
      }
    wave_lds_sync();
  }
  return bad;
}

constexpr int MI_LD = 17;               // row stride of a 16x16 inverse block (doubles)
constexpr int MI_SZ = 16 * MI_LD;       // one inverse block

// Mi = L_pp^{-1} for the 16x16 diagonal sub-block p of L (Ls, row-major LDS;
// rdiag[j] = 1/L_jj), stored row-major with row stride MI_LD.  One wavefront:
// lane k (< 16) forms column k by forward substitution; lanes 16.. repeat the
// work of lane & 15 and do not store.
__device__ __forceinline__ void invert_diag16_wave(const double* Ls, const double* rdiag, double* Mi, int p,
                                                   int lane) {
  const int k = lane & 15;
  double m[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) m[i] = (i == k) ? 1.0 : 0.0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    m[j] *= rdiag[16 * p + j];
#pragma unroll
    for (int i = j + 1; i < 16; ++i) m[i] -= Ls[(16 * p + i) * TS_LD + 16 * p + j] * m[j];
  }
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < 16; ++i) Mi[i * MI_LD + k] = m[i];
  }
}

// Block p of the solve of 16 rows of X (LDS, row-major) in place,
// X <- X L^{-T}, with one wavefront on FP64 MFMA (v_mfma_f64_16x16x4, one
// 16x16 accumulator): X_p -= X_{<p} L_{p,<p}^T, then X_p <- X_p (L_pp^{-1})^T
// with the inverse of the diagonal sub-block (Mi + p * MI_SZ, from
// invert_diag16_wave).  Ls = L (row-major, LDS).  Blocks must be applied in
// order p = 0..3; columns 16p.. of L's rows 16p..16p+15 must be final.
__device__ __forceinline__ void solve_block16_wave(const double* Ls, const double* Mi, double* X, int p,
                                                   int lane) {
  const int row = lane & 15;
  if (p > 0) {
    // X[:, 16p..16p+16) -= X[:, 0..16p) L[16p..16p+16, 0..16p)^T
    doublex4 acc = (doublex4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kb = 0; kb < 16 * p; kb += 4) {
      const int kk = kb + (lane >> 4);
      const double av = X[row * TS_LD + kk];
      const double bv = Ls[(16 * p + row) * TS_LD + kk];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
    // element e of lane l is (row l/16 + 4e, col l%16)
#pragma unroll
    for (int e = 0; e < 4; ++e) X[((lane >> 4) + 4 * e) * TS_LD + 16 * p + row] -= acc[e];
    wave_lds_sync();
  }
  // X_p <- X_p Mi_p^T: operand B[k][n] = Mi_p[n][k]
  doublex4 acc = (doublex4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kb = 0; kb < 16; kb += 4) {
    const int kk = kb + (lane >> 4);
    const double av = X[row * TS_LD + 16 * p + kk];
    const double bv = Mi[p * MI_SZ + row * MI_LD + kk];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  wave_lds_sync();  // every lane's operand reads of X_p precede the writes
#pragma unroll
  for (int e = 0; e < 4; ++e) X[((lane >> 4) + 4 * e) * TS_LD + 16 * p + row] = acc[e];
  wave_lds_sync();
}

// Wavefronts 1-3 while wavefront 0 factors the diagonal block with
// done = pflags[0..3]: wavefront p+1 inverts sub-block p (p < 3) of L11 once
// it is published and announces the inverse in pflags[4+p]; every one of them
// then applies block p of the solve to its 16-row strips (w - 1, and 3 for
// wavefront 1).  Block 3 follows once the factor is done; the operations and
// their order are those of applying blocks 0..3 one after another.
__device__ __forceinline__ void solve_during_factor(const double* Ls, const double* rdiag, double* Mi,
                                                    double (*Xs)[PF_ROWS * TS_LD], int* pflags, int w,
                                                    int lane, int32_t* info, int64_t k0) {
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    if (w == p + 1) {
      lds_wait(&pflags[p], info, k0);
      invert_diag16_wave(Ls, rdiag, Mi + p * MI_SZ, p, lane);
      lds_publish(&pflags[4 + p]);
    }
    lds_wait(&pflags[4 + p], info, k0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int s = w - 1 + 3 * t;
      if (s < 4) solve_block16_wave(Ls, Mi, Xs[s], p, lane);
    }
  }
}

// Factor the diagonal block A[k0:k0+nb, k0:k0+nb] in place (lower triangle)
// with one wavefront; used for the last block, which has no rows below it.
// The factor (64x64 row-major, identity padding past nb, zeros above the
// diagonal) is also written to `slot`, where the inversion kernel reads it.
// A non-positive or non-finite pivot sets *info (first failing column + 1).
__global__ void __launch_bounds__(64) potf2_kernel(double* __restrict__ A, int64_t lda, int64_t n,
                                                   int64_t k0, double* __restrict__ slot,
                                                   int32_t* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double Ab[CH_NB * TS_LD];   // the block, row-major
  __shared__ double colbuf[CH_NB];
  __shared__ double yv[CH_NB];
  const int lane = threadIdx.x;
  const int nb = (int)min<int64_t>(CH_NB, n - k0);
  if (nb == CH_NB) {
    // lane = column; all 64 row loads are issued before the first LDS store
    const double* pc = A + k0 * lda + k0 + lane;
    double v[CH_NB];
#pragma unroll
    for (int r = 0; r < CH_NB; ++r) v[r] = pc[r * lda];
#pragma unroll
    for (int r = 0; r < CH_NB; ++r) Ab[r * TS_LD + lane] = lane <= r ? v[r] : 0.0;
  } else {
    for (int idx = lane; idx < CH_NB * CH_NB; idx += 64) {
      const int r = idx >> 6, c = idx & 63;
      double v = (r == c) ? 1.0 : 0.0;
      if (r < nb && c < nb && c <= r) v = A[(k0 + r) * lda + k0 + c];
      Ab[r * TS_LD + c] = v;
    }
  }
  wave_lds_sync();
  const int bad = factor_block_wave(Ab, colbuf, yv, lane, k0);
#pragma unroll 16
  for (int r = 0; r < nb; ++r) {
    const double v = Ab[r * TS_LD + lane];
    if (lane < nb && lane <= r) A[(k0 + r) * lda + k0 + lane] = v;
  }
#pragma unroll
  for (int r = 0; r < CH_NB; ++r) slot[r * CH_NB + lane] = Ab[r * TS_LD + lane];
  if (lane == 0 && bad) atomicCAS(info, 0, bad);
}

// One 64-column step of the factorisation for a block with rows below it:
// factor the diagonal block A[k0:k0+64, k0:k0+64] and solve the rows below,
// L21 = A21 L11^{-T}, in place.  Every workgroup (four wavefronts of 16 rows,
// 64 rows) reads the diagonal block and factors it in LDS itself (wavefront
// 0), so the solve needs no second launch and no exchange through global
// memory.  Workgroup 0 writes L11 to `slot` (64x64 row-major) and reports a
// failed pivot in *info; L11 reaches A's diagonal block in the inversion
// kernel at the end, since a workgroup that starts late still reads the
// unfactored block from A.  All global loads of a workgroup are issued before
// its first LDS store.  Rows past n re-read the last row and are not stored.

__global__ void __launch_bounds__(256) panel_factor_kernel(double* __restrict__ A, int64_t lda, int64_t n,
                                                           int64_t k0, double* __restrict__ slot,
                                                           int32_t* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double Ls[CH_NB * TS_LD];           // the diagonal block, then L11 (row-major)
  __shared__ double colbuf[CH_NB];
  __shared__ double rdiag[CH_NB];
  __shared__ double Xs[4][PF_ROWS * TS_LD];      // 16 rows per wavefront
  __shared__ double Mi[4 * MI_SZ];               // inverses of L11's diagonal 16x16 blocks
  __shared__ int pflags[8];                      // sub-panels factored, inverses ready
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  double* X = Xs[w];
  if (tid < 8) pflags[tid] = 0;                  // ordered by the barrier below
  const int64_t r0 = k0 + CH_NB + (int64_t)blockIdx.x * (4 * PF_ROWS) + PF_ROWS * w;
  {
    // wavefront w reads rows 16w..16w+15 of the diagonal block and its own
    // 16 rows, lanes along the columns
    double vd[PF_ROWS], v[PF_ROWS];
    const double* pd = A + (k0 + PF_ROWS * w) * lda + k0 + lane;
#pragma unroll
    for (int r = 0; r < PF_ROWS; ++r) vd[r] = pd[r * lda];
#pragma unroll
    for (int r = 0; r < PF_ROWS; ++r) v[r] = A[min<int64_t>(r0 + r, n - 1) * lda + k0 + lane];
#pragma unroll
    for (int r = 0; r < PF_ROWS; ++r) {
      const int R = PF_ROWS * w + r;
      Ls[R * TS_LD + lane] = lane <= R ? vd[r] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < PF_ROWS; ++r) X[r * TS_LD + lane] = v[r];
  }
  __syncthreads();
  // wavefront 0 factors the diagonal block while wavefronts 1-3 solve blocks
  // 0..2 of the four strips as its sub-panels are published
  if (w == 0) {
    const int bad = factor_block_wave(Ls, colbuf, rdiag, lane, k0, pflags);
    if (blockIdx.x == 0 && lane == 0 && bad) atomicCAS(info, 0, bad);
  } else {
    solve_during_factor(Ls, rdiag, Mi, Xs, pflags, w, lane, info, k0);
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    // wavefront w writes rows 16w..16w+15 of L11 (zeros above the diagonal)
    // to the slot; A's diagonal block is left unchanged here because other
    // workgroups of this launch may still be reading it
#pragma unroll
    for (int r = 0; r < PF_ROWS; ++r) {
      const int R = PF_ROWS * w + r;
      slot[R * CH_NB + lane] = Ls[R * TS_LD + lane];
    }
  }
  // the last diagonal sub-block's inverse, then block 3 of each strip
  if (w == 0) invert_diag16_wave(Ls, rdiag, Mi + 3 * MI_SZ, 3, lane);
  __syncthreads();
  solve_block16_wave(Ls, Mi, X, 3, lane);
#pragma unroll
  for (int r = 0; r < PF_ROWS; ++r) {
    if (r0 + r < n) A[(r0 + r) * lda + k0 + lane] = X[r * TS_LD + lane];
  }
}

// Update of the panel columns to the right of a 64-column step:
// C[i][j] -= sum_{k<64} P[i][k] P[j][k] for i < rest, j < cols, where
// P = A[k0+64.., k0..k0+64) holds the solved rows and C = A[k0+64.., k0+64..).
// One 64x64 tile per workgroup.  The operand tiles and the C tile are read
// with every load in flight (one memory latency per tile); the product runs on
// 4 waves of 32x32 (2x2 v_mfma_f64_16x16x4) out of LDS in two K halves.
// Rows and columns past the edge re-read the last valid one and are not stored.
constexpr int PU_LD = 33;  // doubles per row of a 64x32 LDS half tile

// One 64x64 tile of C - Pa Pb^T with K = 64: rows i0.. of Pa (clamped to ra
// rows) against rows j0.. of Pb (clamped to rb rows) as the columns, C read at
// C[(i0 + r) * lda + j0 + c] with the same clamps.  out[x][y][e] is element
// (wm + 16x + l/16 + 4e, wn + 16y + l%16) of the tile, wave (wm, wn) of the
// 2 x 2 wave grid.  All threads of the workgroup call it.
__device__ __forceinline__ void pu_tile(const double* __restrict__ Pa, const double* __restrict__ Pb,
                                        const double* __restrict__ C, int64_t lda, int64_t i0, int64_t j0,
                                        int64_t ra, int64_t rb, double* sA, double* sB,
                                        double out[2][2][4]) {
  const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  // operand loads: wave w reads rows w, w+4, ... with lanes along k
  double va[16], vb[16], vc[2][2][4];
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const int r = wid + 4 * it;
    va[it] = Pa[min<int64_t>(i0 + r, ra - 1) * lda + lane];
    vb[it] = Pb[min<int64_t>(j0 + r, rb - 1) * lda + lane];
  }
  // C in the MFMA output layout: element e of lane l in a 16x16 block is
  // (row l/16 + 4e, col l%16)
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t gi = min<int64_t>(i0 + wm + x * 16 + (lane >> 4) + 4 * e, ra - 1);
        const int64_t gj = min<int64_t>(j0 + wn + y * 16 + (lane & 15), rb - 1);
        vc[x][y][e] = C[gi * lda + gj];
      }
  doublex4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = (doublex4){0.0, 0.0, 0.0, 0.0};
  __syncthreads();  // the staging buffers may still be read by an earlier tile
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h) __syncthreads();
    if ((lane >> 5) == h) {
#pragma unroll
      for (int it = 0; it < 16; ++it) {
        const int r = wid + 4 * it;
        sA[r * PU_LD + (lane & 31)] = va[it];
        sB[r * PU_LD + (lane & 31)] = vb[it];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 32; ks += 4) {
      const int kq = ks + (lane >> 4);
      double av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) av[x] = sA[(wm + x * 16 + (lane & 15)) * PU_LD + kq];
#pragma unroll
      for (int y = 0; y < 2; ++y) bv[y] = sB[(wn + y * 16 + (lane & 15)) * PU_LD + kq];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) out[x][y][e] = vc[x][y][e] - acc[x][y][e];
}

// Block (bi, bj) of the previous step's rank-64 update of the diagonal block,
// C - P P^T with K = 64, from P's rows staged in Pd (64 rows, stride TS_LD)
// and the block's C values cv in the accumulator layout; written to Ls with
// zeros above the diagonal.  The sums run in ascending K steps of 4, the
// order of pu_tile.  One wavefront.
__device__ __forceinline__ void diag_block_update(const double* Pd, double* Ls, int bi, int bj,
                                                  const double cv[4], int lane) {
  doublex4 acc = (doublex4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kb = 0; kb < CH_NB; kb += 4) {
    const int kk = kb + (lane >> 4);
    const double av = Pd[(16 * bi + (lane & 15)) * TS_LD + kk];
    const double bv = Pd[(16 * bj + (lane & 15)) * TS_LD + kk];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  // element e of lane l is (row l/16 + 4e, col l%16) of the block
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int R = 16 * bi + (lane >> 4) + 4 * e, c = 16 * bj + (lane & 15);
    Ls[R * TS_LD + c] = c <= R ? cv[e] - acc[e] : 0.0;
  }
}

// The lower blocks of the diagonal block right of its first 16 columns,
// (1,1) (2,1) | (2,2) (3,1) | (3,2) (3,3), two for each of wavefronts 1-3.
__device__ inline int rest_bi(int w, int t) { return w == 1 ? 1 + t : (w == 2 ? 2 + t : 3); }
__device__ inline int rest_bj(int w, int t) { return w == 1 ? 1 : (w == 2 ? 2 - t : 2 + t); }

// The previous step's update of the diagonal block (P rows 0..63, which must
// all be valid rows) for its first 16 columns, which the factorisation needs
// first: wavefront w updates block (w, 0).  P's rows are staged in Pd, the
// zeros above the diagonal are written, and wavefronts 1-3 also load the C
// values of their two further blocks into cr (applied by diag_block_update
// during the factorisation of the first sub-panel).  All threads of the
// workgroup call it; it ends with a barrier.
__device__ __forceinline__ void diag_col0_update(const double* __restrict__ P, const double* __restrict__ C,
                                                 int64_t lda, double* Pd, double* Ls, double cr[2][4]) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  double vb[16], c0[4];
#pragma unroll
  for (int it = 0; it < 16; ++it) vb[it] = P[(int64_t)(w + 4 * it) * lda + lane];
#pragma unroll
  for (int e = 0; e < 4; ++e) c0[e] = C[(int64_t)(16 * w + (lane >> 4) + 4 * e) * lda + (lane & 15)];
  if (w > 0) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int bi = rest_bi(w, t), bj = rest_bj(w, t);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        cr[t][e] = C[(int64_t)(16 * bi + (lane >> 4) + 4 * e) * lda + 16 * bj + (lane & 15)];
    }
  }
#pragma unroll
  for (int it = 0; it < 16; ++it) Pd[(w + 4 * it) * TS_LD + lane] = vb[it];
  // rows 16w..16w+15 right of their diagonal sub-block hold zeros
  if (lane >= 16 * (w + 1)) {
#pragma unroll
    for (int r = 0; r < 16; ++r) Ls[(16 * w + r) * TS_LD + lane] = 0.0;
  }
  __syncthreads();
  diag_block_update(Pd, Ls, w, 0, c0, lane);
  __syncthreads();
}

// Rows 16s..16s+15 of the tile of rows i0.. of C - P P^T (K = 64; P rows
// i0.. clamped to ra rows, against P rows 0..63 staged in Pd), for one
// wavefront: out[y][e] is element ((l/16) + 4e, 16y + l%16) of the 16-row
// strip, lane l.  The A operand is read from global memory in the MFMA's
// layout; the sums run in the order of pu_tile.
__device__ __forceinline__ void own_strip_update(const double* __restrict__ P, const double* __restrict__ C,
                                                 int64_t lda, int64_t i0, int64_t ra, int s, const double* Pd,
                                                 double out[4][4]) {
  const int lane = threadIdx.x & 63;
  double av[16], cv[4][4];
  const double* pa = P + min<int64_t>(i0 + 16 * s + (lane & 15), ra - 1) * lda + (lane >> 4);
#pragma unroll
  for (int q = 0; q < 16; ++q) av[q] = pa[4 * q];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double* pc = C + min<int64_t>(i0 + 16 * s + (lane >> 4) + 4 * e, ra - 1) * lda + (lane & 15);
#pragma unroll
    for (int y = 0; y < 4; ++y) cv[y][e] = pc[16 * y];
  }
#pragma unroll
  for (int y = 0; y < 4; ++y) {
    doublex4 acc = (doublex4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const double bv = Pd[(16 * y + (lane & 15)) * TS_LD + 4 * q + (lane >> 4)];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) out[y][e] = cv[y][e] - acc[e];
  }
}

// C[i][j] -= sum_{k<64} Pa[i][k] Pb[j][k] over a grid of 64x64 tiles, stored
// for i < rest, j < cols.
__device__ __forceinline__ void pu_store_tile(const double* Pa, const double* Pb, double* C, int64_t lda,
                                              int64_t tile, int64_t tiles_n, int64_t rest, int64_t cols,
                                              double* sA, double* sB) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  const int64_t i0 = (tile / tiles_n) * 64, j0 = (tile % tiles_n) * 64;
  double out[2][2][4];
  pu_tile(Pa, Pb, C, lda, i0, j0, rest, cols, sA, sB, out);
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t gi = i0 + wm + x * 16 + (lane >> 4) + 4 * e;
        const int64_t gj = j0 + wn + y * 16 + (lane & 15);
        if (gi < rest && gj < cols) C[gi * lda + gj] = out[x][y][e];
      }
}

__global__ void __launch_bounds__(256) panel_update_kernel(double* __restrict__ A, int64_t lda, int64_t k0,
                                                           int64_t rest, int64_t cols, int64_t tiles_n,
                                                           int64_t n_tiles, const int32_t* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double sA[CH_NB * PU_LD];
  __shared__ double sB[CH_NB * PU_LD];
  const int64_t bid = xcd_remap(blockIdx.x, gridDim.x);
  if (bid >= n_tiles) return;
  const double* P = A + (k0 + CH_NB) * lda + k0;
  double* C = A + (k0 + CH_NB) * lda + k0 + CH_NB;
  pu_store_tile(P, P, C, lda, bid, tiles_n, rest, cols, sA, sB);
}

// One 64-column step k0 of a panel after its first, in one launch with two
// roles.  Workgroups [0, n_factor) work as panel_factor_kernel, except that
// the previous step's rank-64 update of this step's columns is applied here:
// to the diagonal block's first 16 columns by all four wavefronts before the
// factorisation, to its further lower blocks and to the workgroup's own 64
// rows by wavefronts 1-3 while wavefront 0 factors (same sums, in the same
// order, as the panel update's tiles).  The other workgroups apply the previous step's update to
// the panel's columns right of this step (panel_update_kernel's tiles, shifted
// by 64 columns).  The roles write disjoint columns and read only what earlier
// launches wrote, so the panel update no longer sits between two steps.
__global__ void __launch_bounds__(256, 2) panel_step_kernel(double* __restrict__ A, int64_t lda, int64_t n,
                                                         int64_t k0, int64_t n_factor, int64_t cols_b,
                                                         int64_t tiles_n_b, double* __restrict__ slot,
                                                         int32_t* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double Ls[CH_NB * TS_LD];           // the diagonal block, then L11 (row-major)
  __shared__ double colbuf[CH_NB];
  __shared__ double rdiag[CH_NB];
  __shared__ double Xs[4][PF_ROWS * TS_LD];      // 16 rows per wavefront
  __shared__ double Mi[4 * MI_SZ];               // inverses of L11's diagonal 16x16 blocks
  // LDS flags: [0..3] sub-panels factored, [4..6] their inverses, [8..10]
  // the diagonal block's further blocks updated and [12..14] the own strips
  // updated (wavefronts 1-3)
  __shared__ int pflags[16];
  double* sA = Ls;                               // staging of the update tiles (before Ls is filled)
  double* sB = &Xs[0][0];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int64_t kp = k0 - CH_NB;                 // the previous step
  const double* P = A + k0 * lda + kp;           // rows k0.. of the previous step's columns
  const int64_t rest_p = n - k0;                 // rows of P
  if ((int64_t)blockIdx.x >= n_factor) {
    const int64_t tile = (int64_t)blockIdx.x - n_factor;
    // columns k0 + 64 .. of the panel: B operand rows k0 + 64 ..
    pu_store_tile(P, P + CH_NB * lda, A + k0 * lda + k0 + CH_NB, lda, tile, tiles_n_b, rest_p, cols_b, sA, sB);
    return;
  }
  // diagonal block (rows and columns k0 .. k0+63) first: the launch has rows
  // below this step, so rest_p > 64 and P rows 0..63 are valid.  P rows
  // 0..63 are staged in the Xs region, which the factorisation does not use.
  double* Pd = &Xs[0][0];
  const double* Ck = A + k0 * lda + k0;
  if (tid < 16) pflags[tid] = 0;                 // ordered by the barriers below
  double cr[2][4];
  diag_col0_update(P, Ck, lda, Pd, Ls, cr);
  double* X = Xs[w];
  const int64_t own0 = CH_NB + (int64_t)blockIdx.x * (4 * PF_ROWS);   // relative to k0
  const int64_t r0 = k0 + own0 + PF_ROWS * w;
  // wavefront 0 factors the diagonal block; meanwhile wavefronts 1-3 update
  // its further lower blocks (awaited before the factor first touches them),
  // update this workgroup's 16-row strips (w - 1, and 3 for wavefront 1) in
  // registers, store them once none of them reads Pd any more, and solve
  // their blocks 0..2 as the factor publishes its sub-panels
  if (w == 0) {
    const int bad = factor_block_wave(Ls, colbuf, rdiag, lane, k0, pflags, &pflags[8], info);
    if (blockIdx.x == 0 && lane == 0 && bad) atomicCAS(info, 0, bad);
  } else {
#pragma unroll
    for (int t = 0; t < 2; ++t) diag_block_update(Pd, Ls, rest_bi(w, t), rest_bj(w, t), cr[t], lane);
    lds_publish(&pflags[7 + w]);
    double ow[2][4][4];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int s = w - 1 + 3 * t;
      if (s < 4) own_strip_update(P, Ck, lda, own0, rest_p, s, Pd, ow[t]);
    }
    lds_publish(&pflags[11 + w]);
#pragma unroll
    for (int i = 12; i < 15; ++i) lds_wait(&pflags[i], info, k0);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int s = w - 1 + 3 * t;
      if (s < 4) {
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
          for (int e = 0; e < 4; ++e) Xs[s][((lane >> 4) + 4 * e) * TS_LD + 16 * y + (lane & 15)] = ow[t][y][e];
      }
    }
    wave_lds_sync();
    solve_during_factor(Ls, rdiag, Mi, Xs, pflags, w, lane, info, k0);
  }
  __syncthreads();  // the factor and blocks 0..2 of every strip are done
  if (blockIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < PF_ROWS; ++r) {
      const int R = PF_ROWS * w + r;
      slot[R * CH_NB + lane] = Ls[R * TS_LD + lane];
    }
  }
  if (w == 0) invert_diag16_wave(Ls, rdiag, Mi + 3 * MI_SZ, 3, lane);
  __syncthreads();
  solve_block16_wave(Ls, Mi, X, 3, lane);
#pragma unroll
  for (int r = 0; r < PF_ROWS; ++r) {
    if (r0 + r < n) A[(r0 + r) * lda + k0 + lane] = X[r * TS_LD + lane];
  }
}

// inv_b = L_bb^{-1} for every 64x64 diagonal block b of the factor; one
// wavefront per block.  The factored block is read from its slot in `inv`
// (row-major, identity padding, zeros above the diagonal), copied into A's
// diagonal block, and replaced by its inverse (row-major): lane k computes
// column k by substitution with the block's columns broadcast from LDS.
__global__ void __launch_bounds__(64) trtri_blocks_kernel(double* __restrict__ A, int64_t lda, int64_t n,
                                                          double* __restrict__ inv,
                                                          const int32_t* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double LT[CH_NB * CH_LT];
  __shared__ double rdiag[CH_NB];
  const int k = threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.x * CH_NB;
  const int nb = (int)min<int64_t>(CH_NB, n - k0);
  double* blk = inv + (int64_t)blockIdx.x * CH_NB * CH_NB;
  {
    // lane = column; all row loads in flight before the stores
    double v[CH_NB];
#pragma unroll
    for (int r = 0; r < CH_NB; ++r) v[r] = blk[r * CH_NB + k];
#pragma unroll
    for (int r = 0; r < CH_NB; ++r) LT[k * CH_LT + r] = v[r];
#pragma unroll
    for (int r = 0; r < CH_NB; ++r)
      if (r < nb && k < nb && k <= r) A[(k0 + r) * lda + k0 + k] = v[r];
  }
  wave_lds_sync();
  rdiag[k] = 1.0 / LT[k * CH_LT + k];
  wave_lds_sync();
  double x[CH_NB];
#pragma unroll
  for (int r = 0; r < CH_NB; ++r) x[r] = (r == k) ? 1.0 : 0.0;
#pragma unroll
  for (int m = 0; m < CH_NB; ++m) {
    const double xm = x[m] * rdiag[m];
    x[m] = xm;
#pragma unroll
    for (int r = m + 1; r < CH_NB; ++r) x[r] -= xm * LT[m * CH_LT + r];
  }
#pragma unroll
  for (int r = 0; r < CH_NB; ++r) blk[r * CH_NB + k] = x[r];
}

// Zero n 32-bit words (a fill kernel rather than a memset node, so that a
// captured graph replays it correctly).
__global__ void zero_u32_kernel(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    p[k] = 0u;
}

// ---------------------------------------------------------------------------
// Triangular solves with the blocked factor, one launch each.  Workgroup k
// owns the 64-row block k; it waits (bounded spin in global memory) for every
// block it depends on, and only ever on blocks with a smaller workgroup id,
// which the hardware dispatched earlier, so the wait always ends.  A wait
// that exceeds the spin bound sets *err and the workgroup proceeds (the
// result is then invalid and reported as such by the caller).
//
// The solved entries are their own flags: the output vector is filled with
// kUnsolved before the launch, each entry is written once, as a 64-bit
// atomic store, and a reader polls the entries it needs with 64-bit atomic
// loads until none holds kUnsolved.  One memory round trip per dependency,
// and no ordering between different entries is needed.  kUnsolved is a
// signalling NaN, which arithmetic never returns; a value with its bit
// pattern would be published as the default NaN all the same.
// ---------------------------------------------------------------------------
constexpr int64_t kSpinLimit = 1LL << 26;
constexpr uint64_t kUnsolved = 0x7FF4D15EA5E0F00DULL;
constexpr uint64_t kDefaultNaN = 0x7FF8000000000000ULL;

// The entry *p (lanes with active), polled with relaxed agent-scope loads
// (they read the L2 copy); inactive lanes get 0.  Wave-uniform loop.
__device__ inline double poll_entry(const double* p, bool active, int32_t* err) {
  uint64_t u = 0;
  int64_t spins = 0;
  while (true) {
    if (active) u = __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!__any(active && u == kUnsolved)) break;
    __builtin_amdgcn_s_sleep(1);
    if (++spins > kSpinLimit) {
      if (lane_id() == 0) atomicOr(err, 1);
      break;
    }
  }
  return active ? __longlong_as_double((long long)u) : 0.0;
}

__device__ inline void publish_entry(double* p, double v) {
  uint64_t u = (uint64_t)__double_as_longlong(v);
  if (u == kUnsolved) u = kDefaultNaN;
  __hip_atomic_store((uint64_t*)p, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void fill_unsolved_kernel(uint64_t* __restrict__ p, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    p[k] = kUnsolved;
}

// forward: y_k = inv_kk (b_k - sum_{j<k} L_kj y_j); block kb = blockIdx.x.
// Each wave walks its share of the blocks j < k from the oldest (finished
// first) to the newest, so only the last tile waits on the chain; the tiles
// of L and the inverse block are loaded before any wait.
__global__ void __launch_bounds__(256) trsv_fwd_lookback(const double* __restrict__ A, int64_t lda, int64_t n,
                                                         const double* __restrict__ inv,
                                                         const double* __restrict__ b, double* y,
                                                         const int32_t* __restrict__ info, int32_t* err) {
  if (*info != 0) return;
  __shared__ double part[4][CH_NB];
  __shared__ double t[CH_NB];
  __shared__ double xs[4][CH_NB];
  __shared__ double Xs[CH_NB * (CH_NB + 1)];
  const int kb = blockIdx.x;
  const int64_t k0 = (int64_t)kb * CH_NB;
  const int nb = (int)min<int64_t>(CH_NB, n - k0);
  const int wid = threadIdx.x / 64, r = lane_id();
  const double* invk = inv + (int64_t)kb * CH_NB * CH_NB;
#pragma unroll
  for (int it = 0; it < CH_NB * CH_NB / 256; ++it) {
    const int idx = threadIdx.x + it * 256;
    Xs[(idx >> 6) * (CH_NB + 1) + (idx & 63)] = invk[idx];
  }
  const double bk = (wid == 0 && r < nb) ? b[k0 + r] : 0.0;
  // rows past nb re-read the last row; their sums are never used
  const double* arow = A + (k0 + min(r, nb - 1)) * lda;
  // four interleaved partial sums: the dependent chain of each 64-term dot
  // product is 16 FMAs instead of 64 (it is on the critical path of the solve)
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int jb = wid; jb < kb; jb += 4) {
    const int64_t j0 = (int64_t)jb * CH_NB;
    double Lrow[CH_NB];
#pragma unroll
    for (int c = 0; c < CH_NB; ++c) Lrow[c] = arow[j0 + c];
    // block jb < kb is a full block: y[j0 + lane] is in bounds
    xs[wid][r] = poll_entry(y + j0 + r, true, err);
    wave_lds_sync();
#pragma unroll
    for (int c = 0; c < CH_NB; ++c) acc[c & 3] += Lrow[c] * xs[wid][c];
    wave_lds_sync();
  }
  part[wid][r] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (wid == 0) {
    t[r] = (r < nb) ? bk - (part[0][r] + part[1][r] + part[2][r] + part[3][r]) : 0.0;
    wave_lds_sync();
    // inv_kk is lower triangular (zeros above the diagonal)
    double sp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < CH_NB; ++c) sp[c & 3] += Xs[r * (CH_NB + 1) + c] * t[c];
    const double s = (sp[0] + sp[1]) + (sp[2] + sp[3]);
    if (r < nb) publish_entry(y + k0 + r, s);
  }
}

// backward: x_k = inv_kk^T (y_k - sum_{j>k} L_jk^T x_j); workgroup w owns block
// kb = nblk - 1 - w so that dependencies point to smaller workgroup ids.  Each
// wave walks its blocks j > k from the farthest (finished first) down to k+1.
__global__ void __launch_bounds__(256) trsv_bwd_lookback(const double* __restrict__ A, int64_t lda, int64_t n,
                                                         const double* __restrict__ inv,
                                                         const double* __restrict__ y, double* x,
                                                         const int32_t* __restrict__ info, int32_t* err) {
  if (*info != 0) return;
  __shared__ double part[4][CH_NB];
  __shared__ double t[CH_NB];
  __shared__ double xs[4][CH_NB];
  __shared__ double Xs[CH_NB * (CH_NB + 1)];
  const int nblk = gridDim.x;
  const int kb = nblk - 1 - blockIdx.x;
  const int64_t k0 = (int64_t)kb * CH_NB;
  const int nb = (int)min<int64_t>(CH_NB, n - k0);
  const int wid = threadIdx.x / 64, r = lane_id();
  const double* invk = inv + (int64_t)kb * CH_NB * CH_NB;
#pragma unroll
  for (int it = 0; it < CH_NB * CH_NB / 256; ++it) {
    const int idx = threadIdx.x + it * 256;
    Xs[(idx >> 6) * (CH_NB + 1) + (idx & 63)] = invk[idx];
  }
  const double yk = (wid == 0 && r < nb) ? y[k0 + r] : 0.0;
  // columns past nb re-read the last column; their sums are never used
  const int64_t col = k0 + min(r, nb - 1);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};   // interleaved partial sums, as in the forward solve
  for (int jb = nblk - 1 - wid; jb > kb; jb -= 4) {
    const int64_t j0 = (int64_t)jb * CH_NB;
    const int mb = (int)min<int64_t>(CH_NB, n - j0);
    // column r of the tile L[j0:j0+mb, k0:k0+nb] (coalesced over r); rows
    // past mb re-read the last row and meet a zero entry of x below
    double Lcol[CH_NB];
#pragma unroll
    for (int c = 0; c < CH_NB; ++c) Lcol[c] = A[(j0 + min(c, mb - 1)) * lda + col];
    xs[wid][r] = poll_entry(x + j0 + r, r < mb, err);
    wave_lds_sync();
#pragma unroll
    for (int c = 0; c < CH_NB; ++c) acc[c & 3] += Lcol[c] * xs[wid][c];
    wave_lds_sync();
  }
  part[wid][r] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  if (wid == 0) {
    t[r] = (r < nb) ? yk - (part[0][r] + part[1][r] + part[2][r] + part[3][r]) : 0.0;
    wave_lds_sync();
    // column r of inv_kk (zeros above the diagonal)
    double sp[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < CH_NB; ++c) sp[c & 3] += Xs[c * (CH_NB + 1) + r] * t[c];
    const double s = (sp[0] + sp[1]) + (sp[2] + sp[3]);
    if (r < nb) publish_entry(x + k0 + r, s);
  }
}

// A failed factorisation leaves NaN in the solution, as LAPACK-based
// cho_solve does on a non-positive-definite matrix.
__global__ void nan_on_info_kernel(double* __restrict__ x, int64_t n, const int32_t* __restrict__ info,
                                   const int32_t* __restrict__ err) {
  if (*info == 0 && *err == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = __longlong_as_double(0x7FF8000000000000LL);
}

// ---------------------------------------------------------------------------
// Vector helpers
// ---------------------------------------------------------------------------
// y = alpha op(A) x + beta y.  trans 0: row dot products (one wave per row);
// trans 1: column sums (thread per column); trans 2: symmetric from lower.
__global__ void __launch_bounds__(256) gemv_n_kernel(const double* __restrict__ A, int64_t lda, int64_t m,
                                                     int64_t n, const double* __restrict__ x, double alpha,
                                                     double beta, double* __restrict__ y, int sym_lower) {
  const int wid = threadIdx.x / 64, lane = lane_id();
  for (int64_t r = (int64_t)blockIdx.x * 4 + wid; r < m; r += (int64_t)gridDim.x * 4) {
    double s = 0.0;
    if (!sym_lower) {
      for (int64_t c = lane; c < n; c += 64) s += A[r * lda + c] * x[c];
    } else {
      for (int64_t c = lane; c <= r; c += 64) s += A[r * lda + c] * x[c];
      for (int64_t c = r + 1 + lane; c < n; c += 64) s += A[c * lda + r] * x[c];
    }
    s = wave_sum(s);
    if (lane == 0) y[r] = alpha * s + (beta != 0.0 ? beta * y[r] : 0.0);
  }
}

__global__ void __launch_bounds__(256) gemv_t_partial(const double* __restrict__ A, int64_t lda, int64_t m,
                                                      int64_t n, const double* __restrict__ x,
                                                      double* __restrict__ part, int64_t rows_per) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per;
  const int64_t r1 = min(m, r0 + rows_per);
  if (c >= n) return;
  double s = 0.0;
  for (int64_t r = r0; r < r1; ++r) s += A[r * lda + c] * x[r];
  part[(int64_t)blockIdx.y * n + c] = s;
}

__global__ void gemv_t_reduce(const double* __restrict__ part, int nparts, int64_t n, double alpha,
                              double beta, double* __restrict__ y) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[(int64_t)p * n + c];
    y[c] = alpha * s + (beta != 0.0 ? beta * y[c] : 0.0);
  }
}

__global__ void check_finite_kernel(const double* __restrict__ x, int64_t n, int32_t* __restrict__ flag) {
  int bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    bad |= !isfinite(x[i]);
  if (__any(bad) && lane_id() == 0) atomicOr(flag, 1);
}

template <class T>
__global__ void sgd_kernel(T* __restrict__ p, const double* __restrict__ dp, int64_t n, T lr) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T g = (T)dp[i];
    const T u = -lr * g;
    p[i] = p[i] + u;
  }
}

__global__ void axpy_kernel(int64_t n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

__global__ void centre_scale_kernel(const double* __restrict__ e, int64_t n, const double* __restrict__ mean,
                                    double scale, double* __restrict__ out) {
  const double mu = *mean;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (e[i] - mu) * scale;
}

// Moments of data (C, L) for the blocking statistics: one workgroup.
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int wid = threadIdx.x / 64;
  __syncthreads();
  if (lane_id() == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x / 64;
  for (int w = 0; w < nw; ++w) t += red[w];
  return t;
}

// Mean of the group means and variance around it (two passes); group g of
// size gs starts at element start(g) of data (C, L).
struct GroupSpec {
  int kind;  // 0: chains (size L); 1: blocks of lb within each chain; 2: half chains
  int C, L, lb, nbc, h;
};
__device__ inline int64_t group_count(const GroupSpec& g) {
  return g.kind == 0 ? g.C : (g.kind == 1 ? (int64_t)g.C * g.nbc : 2 * (int64_t)g.C);
}
__device__ inline double group_mean(const double* d, const GroupSpec& g, int64_t q) {
  int64_t start;
  int sz;
  if (g.kind == 0) {
    start = q * g.L;
    sz = g.L;
  } else if (g.kind == 1) {
    start = (q / g.nbc) * g.L + (q % g.nbc) * g.lb;
    sz = g.lb;
  } else {
    start = (q / 2) * g.L + (q % 2) * g.h;
    sz = g.h;
  }
  double a = 0.0;
  for (int l = 0; l < sz; ++l) a += d[start + l];
  return sz > 0 ? a / (double)sz : 0.0;
}
__device__ inline double group_var(const double* d, const GroupSpec& g, double* red) {
  const int64_t cnt = group_count(g);
  double s = 0.0;
  for (int64_t q = threadIdx.x; q < cnt; q += blockDim.x) s += group_mean(d, g, q);
  const double mu = block_sum(s, red) / (double)cnt;
  double v = 0.0;
  for (int64_t q = threadIdx.x; q < cnt; q += blockDim.x) {
    const double x = group_mean(d, g, q) - mu;
    v += x * x;
  }
  return block_sum(v, red) / (double)cnt;
}

__global__ void __launch_bounds__(1024) stats_kernel(const double* __restrict__ d, int C, int L,
                                                     double* __restrict__ out) {
  __shared__ double red[16];
  const int64_t n = (int64_t)C * L;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += d[i];
  const double mean = block_sum(s, red) / (double)n;
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double x = d[i] - mean;
    v += x * x;
  }
  const double var = block_sum(v, red) / (double)n;
  // block size l_block = max(1, L / 32) and floor(L / l_block) blocks per chain;
  // split R-hat halves drop the last element when L is odd
  GroupSpec g;
  g.C = C;
  g.L = L;
  g.lb = max(1, L / 32);
  g.nbc = L / g.lb;
  g.h = L / 2;
  g.kind = 0;
  const double cvar = group_var(d, g, red);
  g.kind = 1;
  const double bvar = group_var(d, g, red);
  g.kind = 2;
  const double hvar = g.h > 0 ? group_var(d, g, red) : 0.0;
  if (threadIdx.x == 0) {
    out[0] = mean;
    out[1] = var;
    out[2] = cvar;
    out[3] = (double)((int64_t)C * g.nbc);
    out[4] = bvar;
    out[5] = hvar;
    out[6] = (double)g.lb;
    out[7] = (double)n;
  }
}

// Copy the lower triangle onto the upper one (row-major, FP64).
__global__ void symmetrize_kernel(double* __restrict__ a, int64_t n, int64_t lda) {
  const int64_t total = n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / n, j = idx % n;
    if (j > i) a[i * lda + j] = a[j * lda + i];
  }
}

// RBM neural tangent kernel on the bf16x3 Gram tiles (called from the NTK
// entry point): Theta from the split tanh values xh + xl (n x K, ldx), S from
// the packed spins.  Returns false when the spins do not fit the epilogue.
bool launch_gram_rbm(const uint16_t* xh, const uint16_t* xl, int64_t n, int64_t K, int64_t ldx,
                     const uint32_t* bits, int nw, int N, const double* v, const double* c, double hb,
                     double ha, double scale, double shift, double offset, double* out, int64_t ldo,
                     hipStream_t s) {
  const int nwp = (nw + 3) / 4 * 4;
  if (nwp > kGramRbmMaxNwp) return false;
  const int64_t T = (n + 2 * kGramRbmWT - 1) / (2 * kGramRbmWT);
  const int64_t tiles = T * (T + 1) / 2;
  GramRbmEpi ep{bits, nw, nwp, N, v, c, hb, ha, scale};
  hipLaunchKernelGGL((gram_bf16x3_kernel<true, kGramRbmWT>), dim3(grid_mult8(tiles)), dim3(256), 0, s, xh, xl,
                     n, K, ldx, shift, offset, out, ldo, tiles, ep);
  return true;
}

// The two-layer network's tangent kernel on the bf16x3 Gram tiles (called from
// the network's NTK entry point): H, G and D from the split h1, g2 and d1
// (hi + lo planes, n rows each), X from the packed spins.  Returns false when
// the spins do not fit the epilogue.
// C = A (Bh + Bl)^T (FP32) for an exact single-plane A; K and ldk multiples
// of 8, operands 16-byte aligned (checked by the caller).
void launch_gemm_exact_a(const uint16_t* a, int64_t ma, const uint16_t* bh, const uint16_t* bl, int64_t nb,
                         int64_t K, int64_t ldk, float* c, int64_t ldc, hipStream_t s) {
  constexpr int TA = 64 * kExactAMIA, TB = 64 * kExactAMIB;
  const int64_t tm = (ma + TA - 1) / TA, tn = (nb + TB - 1) / TB;
  if (tm == 0 || tn == 0) return;
  hipLaunchKernelGGL((gemm_exact_a_kernel<kExactAMIA, kExactAMIB>), dim3(grid_mult8(tm * tn)), dim3(256), 0, s, a,
                     ma, bh, bl, nb, K, ldk, c, ldc, tm, tm * tn);
}

bool launch_gram_ffnn(const uint16_t* h1h, const uint16_t* h1l, int64_t K0, int64_t ld0, const uint16_t* g2h,
                      const uint16_t* g2l, int64_t K1, int64_t ld1, const uint16_t* d1h, const uint16_t* d1l,
                      int64_t K2, int64_t ld2, int64_t n, const uint32_t* bits, int nw, int N, const double* v,
                      const double* c, double hb1, double hb2, double scale, double shift, double offset,
                      double* out, int64_t ldo, hipStream_t s) {
  const int nwp = (nw + 3) / 4 * 4;
  if (nwp > kGramRbmMaxNwp) return false;
  const int64_t T = (n + 2 * kGramRbmWT - 1) / (2 * kGramRbmWT);
  const int64_t tiles = T * (T + 1) / 2;
  GramRbmEpi ep{bits, nw, nwp, N, v, c, hb1, 0.0, scale, g2h, g2l, d1h, d1l, K1, K2, ld1, ld2, hb2};
  hipLaunchKernelGGL((gram_bf16x3_kernel<true, kGramRbmWT, 3>), dim3(grid_mult8(tiles)), dim3(256), 0, s, h1h,
                     h1l, n, K0, ld0, shift, offset, out, ldo, tiles, ep);
  return true;
}

}  // namespace nka

using namespace nka;

extern "C" {

int nka_gram_bf16x3(const uint16_t* x_hi, const uint16_t* x_lo, int64_t n, int64_t k, int64_t ldx,
                    double shift, double offset, double* out, int64_t ldo, void* stream) {
  NKA_CHECK_ARG(n >= 0 && k >= 0 && ldx >= k && ldo >= n, "nka_gram_bf16x3: bad sizes");
  NKA_CHECK_ARG(ldx % 8 == 0 && k % 8 == 0,
                "nka_gram_bf16x3: k and ldx must be multiples of 8 (pad with zeros)");
  NKA_CHECK_ARG(((uintptr_t)x_hi % 16) == 0 && ((uintptr_t)x_lo % 16) == 0,
                "nka_gram_bf16x3: operands must be 16-byte aligned");
  if (n == 0) return NKA_OK;
  const int64_t T = (n + 2 * kGramWT - 1) / (2 * kGramWT);
  const int64_t tiles = T * (T + 1) / 2;
  NKA_CHECK_ARG(tiles < (1LL << 31), "nka_gram_bf16x3: matrix too large");
  hipLaunchKernelGGL((gram_bf16x3_kernel<false, kGramWT>), dim3(grid_mult8(tiles)), dim3(256), 0,
                     (hipStream_t)stream, x_hi, x_lo, n, k, ldx, shift, offset, out, ldo, tiles, GramRbmEpi{});
  NKA_CHECK_LAUNCH("nka_gram_bf16x3");
  return NKA_OK;
}

static int gemm_launch(const double* a, int64_t lda, const double* b, int64_t ldb, double* c,
                       int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta,
                       double shift, int lower, const int32_t* info, hipStream_t s) {
  const int64_t tm = (m + GD_T - 1) / GD_T, tn = (n + GD_T - 1) / GD_T;
  const int64_t tiles = lower ? tm * (tm + 1) / 2 : tm * tn;
  if (tiles == 0) return NKA_OK;
  hipLaunchKernelGGL(gemm_nt_f64_kernel, dim3(grid_mult8(tiles)), dim3(256), 0, s, a, lda, b, ldb, c,
                     ldc, m, n, k, alpha, beta, shift, lower, tn, tiles, info);
  return NKA_OK;
}

// The 128-tile kernel needs K to be a multiple of its K step; other shapes
// take the 64-tile kernel.
// Workgroups of the 128-tile kernel resident at once (two per CU), from the
// device's CU count (queried once per process for the current device).
static int64_t big_gemm_slots() {
  static int64_t slots = 0;
  if (slots == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      slots = 2 * (int64_t)cus;
    else
      slots = 512;
  }
  return slots;
}

// Tail-split workspace of the 128-tile GEMM: at most kMaxSplitUnits raw tiles.
constexpr int64_t kMaxSplitUnits = 1024;

// The 128-tile kernel needs K to be a multiple of its K step; other shapes
// take the 64-tile kernel.  With a workspace, the tiles of the last, partial
// round of the grid are split along K so that the round is filled, and their
// slices are summed into C by a second launch.
static void gemm_big_launch(const double* a, int64_t lda, const double* b, int64_t ldb, double* c,
                            int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta,
                            int lower, const int32_t* info, hipStream_t s, double* work = nullptr,
                            int64_t work_units = 0) {
  if (m <= 0 || n <= 0) return;
  if (k % G2_K != 0) {
    gemm_launch(a, lda, b, ldb, c, ldc, m, n, k, alpha, beta, 0.0, lower, info, s);
    return;
  }
  const int64_t tm = (m + G2_T - 1) / G2_T, tn = (n + G2_T - 1) / G2_T;
  const int64_t tiles = lower ? tm * (tm + 1) / 2 : tm * tn;
  const int64_t slots = big_gemm_slots();
  const int64_t n_main = (tiles / slots) * slots;
  const int64_t tail = tiles - n_main;
  // slices per tail tile: fill the round, keep >= 4 K steps per slice
  int S = 1;
  if (work && tail > 0) {
    const int64_t max_s = std::max<int64_t>(1, k / (4 * G2_K));
    S = (int)std::min<int64_t>(std::min<int64_t>(slots / tail, max_s), work_units / tail);
    if (S < 2) S = 1;
  }
  const int64_t kc = S > 1 ? ((k + S - 1) / S + G2_K - 1) / G2_K * G2_K : k;
  if (S > 1) S = (int)((k + kc - 1) / kc);  // slices actually needed with kc columns each
  const int64_t main_tiles = S > 1 ? n_main : tiles;
  const bool vec = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && lda % 2 == 0 && ldb % 2 == 0;
  if (main_tiles > 0) {
    const unsigned grid = grid_mult8(main_tiles);
    if (vec)
      hipLaunchKernelGGL((gemm_nt_f64_big_kernel<true, false>), dim3(grid), dim3(256), 0, s, a, lda, b, ldb, c,
                         ldc, m, n, k, alpha, beta, lower, tn, main_tiles, (int64_t)0, 1, k, work, info);
    else
      hipLaunchKernelGGL((gemm_nt_f64_big_kernel<false, false>), dim3(grid), dim3(256), 0, s, a, lda, b, ldb,
                         c, ldc, m, n, k, alpha, beta, lower, tn, main_tiles, (int64_t)0, 1, k, work, info);
  }
  if (S > 1) {
    const int64_t su = tail * S;
    const unsigned grid = grid_mult8(su);
    if (vec)
      hipLaunchKernelGGL((gemm_nt_f64_big_kernel<true, true>), dim3(grid), dim3(256), 0, s, a, lda, b, ldb, c,
                         ldc, m, n, k, alpha, beta, lower, tn, su, main_tiles, S, kc, work, info);
    else
      hipLaunchKernelGGL((gemm_nt_f64_big_kernel<false, true>), dim3(grid), dim3(256), 0, s, a, lda, b, ldb,
                         c, ldc, m, n, k, alpha, beta, lower, tn, su, main_tiles, S, kc, work, info);
  }
  if (S > 1)
    hipLaunchKernelGGL(gemm_big_reduce_kernel, dim3((unsigned)(tail * 4)), dim3(256), 0, s, work, c, ldc, m, n,
                       alpha, beta, lower, tn, main_tiles, S, info);
}

int nka_syrk_f64(const double* x, int64_t n, int64_t k, int64_t ldx, double alpha, double beta,
                 double shift, double* c, int64_t ldc, void* stream) {
  NKA_CHECK_ARG(n >= 0 && k >= 0 && ldx >= k && ldc >= n, "nka_syrk_f64: bad sizes");
  gemm_launch(x, ldx, x, ldx, c, ldc, n, n, k, alpha, beta, shift, 1, nullptr, (hipStream_t)stream);
  NKA_CHECK_LAUNCH("nka_syrk_f64");
  return NKA_OK;
}

int nka_gemm_nt_f64(const double* a, int64_t lda, const double* b, int64_t ldb, double* c,
                    int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta,
                    void* stream) {
  NKA_CHECK_ARG(m >= 0 && n >= 0 && k >= 0 && lda >= k && ldb >= k && ldc >= n,
                "nka_gemm_nt_f64: bad sizes");
  if (k == 0) {
    gemm_launch(a, lda, b, ldb, c, ldc, m, n, k, alpha, beta, 0.0, 0, nullptr, (hipStream_t)stream);
  } else {
    // 128x128 tiles when K is a multiple of their K step, else 64x64 tiles
    gemm_big_launch(a, lda, b, ldb, c, ldc, m, n, k, alpha, beta, 0, nullptr, (hipStream_t)stream);
  }
  NKA_CHECK_LAUNCH("nka_gemm_nt_f64");
  return NKA_OK;
}

// Workspace: 64x64 inverses of the diagonal blocks, the intermediate vector
// of the solves, one flag per block and one error word, and the K-slices of
// the trailing products' split tail tiles (up to kMaxSplitUnits raw 128x128
// tiles, fewer for small n).
static int64_t split_units(int64_t n) {
  const int64_t t = (n + G2_T - 1) / G2_T;
  return std::min<int64_t>(kMaxSplitUnits, 2 * t * (t + 1) / 2);
}

int64_t nka_potrf_workspace(int64_t n) {
  if (n <= 0) return 0;
  const int64_t nblk = (n + CH_NB - 1) / CH_NB;
  const int64_t head = nblk * CH_NB * CH_NB * (int64_t)sizeof(double) + (n + 64) * (int64_t)sizeof(double) +
                       (nblk + 64) * (int64_t)sizeof(uint32_t);
  return ((head + 255) / 256) * 256 + split_units(n) * G2_T * G2_T * (int64_t)sizeof(double);
}

static inline double* ws_inv(void* ws) { return (double*)ws; }
static inline double* ws_vec(void* ws, int64_t nblk) {
  return (double*)ws + nblk * CH_NB * CH_NB;
}
static inline uint32_t* ws_flags(void* ws, int64_t n, int64_t nblk) {
  return (uint32_t*)(ws_vec(ws, nblk) + n + 64);
}
static inline double* ws_split(void* ws, int64_t n, int64_t nblk) {
  const int64_t head = nblk * CH_NB * CH_NB * (int64_t)sizeof(double) + (n + 64) * (int64_t)sizeof(double) +
                       (nblk + 64) * (int64_t)sizeof(uint32_t);
  return (double*)((char*)ws + ((head + 255) / 256) * 256);
}

// Right-looking Cholesky with up to three levels of blocking.  An outer
// panel of NB0 columns is factored as panels of NB1 columns, each by 64-column
// steps (diagonal block and panel solve in one launch, then the update of the
// NB1-panel's remaining columns); after each NB1-panel the rest of the outer
// panel receives a rank-NB1 update, and after the outer panel the trailing
// matrix one rank-NB0 update.  A long K keeps the MFMA GEMM near its peak and
// reads and writes the trailing matrix fewer times; the widths trade that
// against the tile-count quantisation of the narrower updates.  The inverses
// of the diagonal blocks, used by the triangular solves, are formed at the
// end.
// The launches of the factorisation.  split / n_split are the K-slices of
// the trailing products' split tail tiles (nullptr / 0: no split).
static void potrf_run(double* a, int64_t n, int64_t lda, void* workspace, double* split, int64_t n_split,
                      int32_t* info, hipStream_t s) {
  double* inv = ws_inv(workspace);
  const int64_t nblk = (n + CH_NB - 1) / CH_NB;
  // panel widths measured best on MI355X: (768, 768) below n = 16384 (at
  // n = 4096, 8384 and 12288: 2.02, 6.64 and 15.1 ms, against 2.06, 6.94 and
  // 15.8 ms with (512, 512)); (4096, 512) from n = 16384 (30.8 ms at 16384,
  // 200 ms at 32768)
  const bool large = n >= 16384;
  const int64_t NB0 = large ? 4096 : 768, NB1 = large ? 512 : 768;
  for (int64_t K0 = 0; K0 < n; K0 += NB0) {
    const int64_t n0 = std::min<int64_t>(NB0, n - K0);
    for (int64_t K1 = K0; K1 < K0 + n0; K1 += NB1) {
      const int64_t n1 = std::min<int64_t>(NB1, K0 + n0 - K1);
      // 64-column steps.  The first step of the panel factors its columns as
      // the trailing updates left them.  A later step runs as one two-role
      // launch: the previous step's update of this step's columns fused into
      // the factorisation, and its update of the columns further right beside
      // it.  The last block of the matrix (no rows below) takes the separate
      // update and the one-wavefront factor.
      for (int64_t k0 = K1; k0 < K1 + n1; k0 += CH_NB) {
        const int64_t nb = std::min<int64_t>(CH_NB, n - k0);
        const int64_t rest = n - k0 - nb;
        double* slot = inv + (k0 / CH_NB) * CH_NB * CH_NB;
        if (rest <= 0) {
          if (k0 > K1) {
            const int64_t kp = k0 - CH_NB, rest_p = n - k0, cols_p = K1 + n1 - k0;
            const int64_t tm = (rest_p + 63) / 64, tn = (cols_p + 63) / 64;
            hipLaunchKernelGGL(panel_update_kernel, dim3(grid_mult8(tm * tn)), dim3(256), 0, s, a, lda, kp,
                               rest_p, cols_p, tn, tm * tn, info);
          }
          hipLaunchKernelGGL(potf2_kernel, dim3(1), dim3(64), 0, s, a, lda, n, k0, slot, info);
          continue;
        }
        // rows below imply nb == 64
        const unsigned grid = (unsigned)((rest + 4 * PF_ROWS - 1) / (4 * PF_ROWS));
        if (k0 == K1) {
          hipLaunchKernelGGL(panel_factor_kernel, dim3(grid), dim3(256), 0, s, a, lda, n, k0, slot, info);
        } else {
          // the previous step's update of the columns right of this step
          const int64_t cols_b = K1 + n1 - (k0 + CH_NB);
          const int64_t tm = (n - k0 + 63) / 64, tn = cols_b > 0 ? (cols_b + 63) / 64 : 0;
          const int64_t tiles_b = cols_b > 0 ? tm * tn : 0;
          hipLaunchKernelGGL(panel_step_kernel, dim3((unsigned)(grid + tiles_b)), dim3(256), 0, s, a, lda, n,
                             k0, (int64_t)grid, cols_b, tn, slot, info);
        }
      }
      // update of the outer panel's columns right of this NB1-column panel
      // (K = NB1): the lower triangle of its square top, then the rows below it
      const int64_t cols1 = K0 + n0 - (K1 + n1);
      if (cols1 > 0) {
        const int64_t rows1 = n - (K1 + n1);
        const double* l = a + (K1 + n1) * lda + K1;
        double* c = a + (K1 + n1) * lda + (K1 + n1);
        gemm_big_launch(l, lda, l, lda, c, lda, cols1, cols1, n1, -1.0, 1.0, 1, info, s, split, n_split);
        if (rows1 > cols1)
          gemm_big_launch(l + cols1 * lda, lda, l, lda, c + cols1 * lda, lda, rows1 - cols1, cols1, n1, -1.0,
                          1.0, 0, info, s, split, n_split);
      }
    }
    // trailing update (K = NB0)
    const int64_t rest2 = n - K0 - n0;
    if (rest2 > 0) {
      const double* l21 = a + (K0 + n0) * lda + K0;
      double* a22 = a + (K0 + n0) * lda + (K0 + n0);
      gemm_big_launch(l21, lda, l21, lda, a22, lda, rest2, rest2, n0, -1.0, 1.0, 1, info, s, split, n_split);
    }
  }
  hipLaunchKernelGGL(trtri_blocks_kernel, dim3((unsigned)nblk), dim3(64), 0, s, a, lda, n, inv, info);
}

int nka_potrf_f64(double* a, int64_t n, int64_t lda, void* workspace, int32_t* info,
                  void* stream) {
  NKA_CHECK_ARG(n >= 0 && lda >= n && a && workspace && info, "nka_potrf_f64: bad arguments");
  if (n == 0) return NKA_OK;
  const int64_t nblk = (n + CH_NB - 1) / CH_NB;
  NKA_CHECK_ARG(nblk < (1LL << 31), "nka_potrf_f64: too large");
  potrf_run(a, n, lda, workspace, ws_split(workspace, n, nblk), split_units(n), info, (hipStream_t)stream);
  NKA_CHECK_LAUNCH("nka_potrf_f64");
  return NKA_OK;
}

// The launches of the two triangular solves.
static void potrs_run(const double* a, int64_t n, int64_t lda, void* ws, double* b, const int32_t* info,
                      hipStream_t s) {
  const int64_t nblk = (n + CH_NB - 1) / CH_NB;
  const double* inv = ws_inv(ws);
  double* y = ws_vec(ws, nblk);
  uint32_t* flags = ws_flags(ws, n, nblk);
  int32_t* err = (int32_t*)(flags + nblk);
  const unsigned gf = (unsigned)std::min<int64_t>((n + 255) / 256, 256);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, (uint32_t*)err, (int64_t)1);
  hipLaunchKernelGGL(fill_unsolved_kernel, dim3(gf), dim3(256), 0, s, (uint64_t*)y, n);
  hipLaunchKernelGGL(trsv_fwd_lookback, dim3((unsigned)nblk), dim3(256), 0, s, a, lda, n, inv, b, y,
                     info, err);
  // b has been read: it now receives x, entry by entry
  hipLaunchKernelGGL(fill_unsolved_kernel, dim3(gf), dim3(256), 0, s, (uint64_t*)b, n);
  hipLaunchKernelGGL(trsv_bwd_lookback, dim3((unsigned)nblk), dim3(256), 0, s, a, lda, n, inv, y, b,
                     info, err);
  hipLaunchKernelGGL(nan_on_info_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 1024)),
                     dim3(256), 0, s, b, n, info, err);
}

int nka_potrs_f64(const double* a, int64_t n, int64_t lda, const void* workspace, double* b,
                  const int32_t* info, void* stream) {
  NKA_CHECK_ARG(n >= 0 && lda >= n && a && workspace && b && info, "nka_potrs_f64: bad arguments");
  if (n == 0) return NKA_OK;
  NKA_CHECK_ARG((n + CH_NB - 1) / CH_NB < (1LL << 31), "nka_potrs_f64: too large");
  potrs_run(a, n, lda, (void*)workspace, b, info, (hipStream_t)stream);
  NKA_CHECK_LAUNCH("nka_potrs_f64");
  return NKA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Factor and solve with no allocation and no caller workspace.  The scratch
// (inverses of the diagonal blocks, the intermediate vector, the flags and
// the error word of the solves) is a static device array of a fixed size;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device.  Trailing products run unsplit here (the split of
// tail tiles would need up to 128 MiB more), so the last bits can differ from
// nka_potrf_f64 + nka_potrs_f64 by the order of the K sums of those tiles.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxN = 32768;
constexpr int64_t kNoAllocNblk = (kNoAllocMaxN + CH_NB - 1) / CH_NB;
constexpr int64_t kNoAllocHead = kNoAllocNblk * CH_NB * CH_NB * (int64_t)sizeof(double) +
                                 (kNoAllocMaxN + 64) * (int64_t)sizeof(double) +
                                 (kNoAllocNblk + 64) * (int64_t)sizeof(uint32_t);
constexpr int64_t kNoAllocBytes = (kNoAllocHead + 255) / 256 * 256;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

// The device address of `scratch` on the current device, looked up once per device.
static void* scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    set_error("nka_cholesky_solve_f64_noalloc: no current HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("nka_cholesky_solve_f64_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// Scale-invariant regularisation of the QGT (netket/optimizer/qgt/qgt_jacobian_common.py:21-60):
// d_k = sqrt(S_kk + offset) from the diagonal of the unregularised S.
__global__ void qgt_diag_scale_kernel(const double* __restrict__ a, int64_t n, int64_t lda, double offset,
                                      double* __restrict__ d) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    d[k] = sqrt(a[k * lda + k] + offset);
}
// mode 0: lower triangle S_ij / (d_i d_j) + shift delta_ij (the rescaled system);
// mode 1: S_kk + shift d_k^2 on the diagonal only (to_dense of the scaled QGT).
__global__ void qgt_rescale_kernel(double* __restrict__ a, int64_t n, int64_t lda, const double* __restrict__ d,
                                   double shift, int mode) {
  if (mode == 1) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
      a[k * lda + k] += shift * d[k] * d[k];
    return;
  }
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const double di = d[i];
    for (int64_t j = threadIdx.x; j <= i; j += blockDim.x) {
      double v = a[i * lda + j] / (di * d[j]);
      if (j == i) v += shift;
      a[i * lda + j] = v;
    }
  }
}
__global__ void mul_div_kernel(const double* __restrict__ x, const double* __restrict__ d, int64_t n, int divide,
                               double* __restrict__ out) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = divide ? x[k] / d[k] : x[k] * d[k];
}

__global__ void select_on_flag_kernel(const int32_t* __restrict__ flag, const double* __restrict__ src,
                                      double* __restrict__ x, int64_t n) {
  if (*flag == 0) return;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    x[k] = src ? src[k] : 0.0;
}

extern "C" {

int64_t nka_cholesky_solve_noalloc_max_n(void) { return kNoAllocMaxN; }

int nka_cholesky_solve_f64_noalloc(double* a, int64_t n, int64_t lda, double* b, int32_t* info,
                                   void* stream) {
  NKA_CHECK_ARG(n >= 0 && lda >= n && info && (n == 0 || (a && b)),
                "nka_cholesky_solve_f64_noalloc: bad arguments");
  NKA_CHECK_ARG(n <= kNoAllocMaxN, "nka_cholesky_solve_f64_noalloc: n = %lld exceeds the maximum %lld of the "
                "static scratch; use nka_potrf_f64 / nka_potrs_f64 with a workspace",
                (long long)n, (long long)kNoAllocMaxN);
  void* ws = scratch_on_current_device();
  if (!ws) return NKA_ERR_HIP;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, (uint32_t*)info, (int64_t)1);
  if (n == 0) {
    NKA_CHECK_LAUNCH("nka_cholesky_solve_f64_noalloc");
    return NKA_OK;
  }
  potrf_run(a, n, lda, ws, nullptr, 0, info, s);
  potrs_run(a, n, lda, ws, b, info, s);
  NKA_CHECK_LAUNCH("nka_cholesky_solve_f64_noalloc");
  return NKA_OK;
}

int nka_qgt_rescale_f64(double* a, int64_t n, int64_t lda, double offset, double shift, int mode,
                        double* scale, void* stream) {
  NKA_CHECK_ARG(n >= 0 && lda >= n && (mode == 0 || mode == 1) && (n == 0 || (a && scale)),
                "nka_qgt_rescale_f64: bad arguments");
  if (n == 0) return NKA_OK;
  hipStream_t s = (hipStream_t)stream;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(qgt_diag_scale_kernel, dim3(g), dim3(256), 0, s, a, n, lda, offset, scale);
  const unsigned g2 = mode == 1 ? g : (unsigned)std::min<int64_t>(n, 8192);
  hipLaunchKernelGGL(qgt_rescale_kernel, dim3(g2), dim3(256), 0, s, a, n, lda, scale, shift, mode);
  NKA_CHECK_LAUNCH("nka_qgt_rescale_f64");
  return NKA_OK;
}

int nka_mul_div_f64(const double* x, const double* d, int64_t n, int divide, double* out, void* stream) {
  NKA_CHECK_ARG(n >= 0 && (n == 0 || (x && d && out)), "nka_mul_div_f64: bad arguments");
  if (n == 0) return NKA_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(mul_div_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, x, d, n, divide, out);
  NKA_CHECK_LAUNCH("nka_mul_div_f64");
  return NKA_OK;
}

int nka_select_on_flag_f64(const int32_t* flag, const double* src, double* x, int64_t n, void* stream) {
  NKA_CHECK_ARG(n >= 0 && flag && (n == 0 || x), "nka_select_on_flag_f64: bad arguments");
  if (n == 0) return NKA_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(select_on_flag_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, flag, src, x, n);
  NKA_CHECK_LAUNCH("nka_select_on_flag_f64");
  return NKA_OK;
}

int nka_gemv_f64(const double* a, int64_t lda, int64_t m, int64_t n, int trans, const double* x,
                 double alpha, double beta, double* y, void* workspace, void* stream) {
  NKA_CHECK_ARG(m >= 0 && n >= 0 && a && x && y, "nka_gemv_f64: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (trans == 0 || trans == 2) {
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((m + 3) / 4, 4096));
    hipLaunchKernelGGL(gemv_n_kernel, dim3(g), dim3(256), 0, s, a, lda, m, n, x, alpha, beta, y,
                       trans == 2 ? 1 : 0);
  } else if (trans == 1) {
    NKA_CHECK_ARG(workspace != nullptr, "nka_gemv_f64: transposed product needs a workspace");
    const int64_t rows_per = 256;
    const int64_t parts = (m + rows_per - 1) / rows_per;
    NKA_CHECK_ARG(parts <= 65535, "nka_gemv_f64: too many rows");
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)std::max<int64_t>(parts, 1));
    hipLaunchKernelGGL(gemv_t_partial, grid, dim3(256), 0, s, a, lda, m, n, x, (double*)workspace,
                       rows_per);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(gemv_t_reduce, dim3(g), dim3(256), 0, s, (const double*)workspace,
                       (int)std::max<int64_t>(parts, 1), n, alpha, beta, y);
  } else {
    set_error("nka_gemv_f64: bad trans %d", trans);
    return NKA_ERR_ARG;
  }
  NKA_CHECK_LAUNCH("nka_gemv_f64");
  return NKA_OK;
}

int nka_check_finite(const double* x, int64_t n, int32_t* flag, void* stream) {
  NKA_CHECK_ARG(n >= 0 && flag && (n == 0 || x), "nka_check_finite: bad arguments");
  if (n == 0) return NKA_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(check_finite_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, x, n, flag);
  NKA_CHECK_LAUNCH("nka_check_finite");
  return NKA_OK;
}

int nka_sgd_update(void* params, int dtype, const double* dp, int64_t n, double lr, void* stream) {
  NKA_CHECK_ARG(n >= 0 && (n == 0 || (params && dp)), "nka_sgd_update: bad arguments");
  if (n == 0) return NKA_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  if (dtype == NKA_F32)
    hipLaunchKernelGGL(sgd_kernel<float>, dim3(g), dim3(256), 0, (hipStream_t)stream,
                       (float*)params, dp, n, (float)lr);
  else if (dtype == NKA_F64)
    hipLaunchKernelGGL(sgd_kernel<double>, dim3(g), dim3(256), 0, (hipStream_t)stream,
                       (double*)params, dp, n, lr);
  else {
    set_error("nka_sgd_update: bad dtype");
    return NKA_ERR_ARG;
  }
  NKA_CHECK_LAUNCH("nka_sgd_update");
  return NKA_OK;
}

int nka_axpy_f64(int64_t n, double alpha, const double* x, double* y, void* stream) {
  NKA_CHECK_ARG(n >= 0 && (n == 0 || (x && y)), "nka_axpy_f64: bad arguments");
  if (n == 0) return NKA_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(axpy_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, n, alpha, x, y);
  NKA_CHECK_LAUNCH("nka_axpy_f64");
  return NKA_OK;
}

int nka_centre_scale(const double* e, int64_t n, const double* mean, double scale, double* out,
                     void* stream) {
  NKA_CHECK_ARG(n >= 0 && (n == 0 || (e && mean && out)), "nka_centre_scale: bad arguments");
  if (n == 0) return NKA_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(centre_scale_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, e, n, mean,
                     scale, out);
  NKA_CHECK_LAUNCH("nka_centre_scale");
  return NKA_OK;
}

int nka_symmetrize_f64(double* a, int64_t n, int64_t lda, void* stream) {
  NKA_CHECK_ARG(n >= 0 && lda >= n && (n == 0 || a), "nka_symmetrize_f64: bad arguments");
  if (n == 0) return NKA_OK;
  const unsigned g = (unsigned)std::min<int64_t>((n * n + 255) / 256, 65536);
  hipLaunchKernelGGL(symmetrize_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, a, n, lda);
  NKA_CHECK_LAUNCH("nka_symmetrize_f64");
  return NKA_OK;
}

int nka_stats_moments(const double* data, int n_chains, int chain_length, double* out8,
                      void* stream) {
  NKA_CHECK_ARG(n_chains > 0 && chain_length > 0 && data && out8, "nka_stats_moments: bad arguments");
  hipLaunchKernelGGL(stats_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, data, n_chains,
                     chain_length, out8);
  NKA_CHECK_LAUNCH("nka_stats_moments");
  return NKA_OK;
}

}  // extern "C"
