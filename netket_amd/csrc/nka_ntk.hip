// Structured neural tangent kernel of the RBM for minSR
// (T = O_L O_L^T + eps I, netket/_src/ngd/srt.py:55-66).
//
// With O[s] = [b? t_s, sigma_s (x) t_s, a? sigma_s] (t = tanh theta) the
// uncentred kernel is
//   (O O^T)_st = (hb + sigma_s.sigma_t) (t_s.t_t) + ha sigma_s.sigma_t
// (hb, ha = 1 when the hidden / visible bias exists), and centring with the
// column means mu gives
//   T_st = scale * [(O O^T)_st - v_s - v_t + c],  v = O mu,  c = |mu|^2,
// scale = 1/Ns.  The inner products have K = M and K = N instead of the
// K = Np of the dense Jacobian product.  FP64 models form both on FP64
// matrix cores; FP32 models form t.t on bf16 matrix cores from the split
// t = hi + lo (three products, FP32 accumulation) and sigma.sigma exactly from
// packed spins in the epilogue (the bf16x3 Gram kernel of nka_linalg.hip).
#include "nka_common.hpp"
#include "netket_amd.h"

namespace nka {

typedef double doublex4 __attribute__((ext_vector_type(4)));

// t64 = (double) t  (ns x M, model dtype -> FP64)
template <class T>
__global__ void to_f64_kernel(const T* __restrict__ t, int64_t total, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (double)t[i];
}

// v_s = hb t_s.mu_b + sum_i sigma_si U_si + ha sigma_s.mu_a, one wavefront per sample;
// block 0 also reduces c = |mu|^2.
__global__ void __launch_bounds__(256) ntk_vec_kernel(const double* __restrict__ t64, const int8_t* __restrict__ sigma,
                                                      const double* __restrict__ U, int64_t ns, int N, int M,
                                                      const double* __restrict__ mean, int64_t b_off,
                                                      int64_t a_off, int64_t np, double* __restrict__ v,
                                                      double* __restrict__ c) {
  const int wid = threadIdx.x / 64, lane = lane_id();
  for (int64_t s = (int64_t)blockIdx.x * 4 + wid; s < ns; s += (int64_t)gridDim.x * 4) {
    double acc = 0.0;
    for (int i = lane; i < N; i += 64) {
      const double sg = sigma[s * N + i] > 0 ? 1.0 : -1.0;
      acc += sg * U[s * N + i];
      if (a_off >= 0) acc += sg * mean[a_off + i];
    }
    if (b_off >= 0)
      for (int j = lane; j < M; j += 64) acc += t64[s * M + j] * mean[b_off + j];
    acc = wave_sum(acc);
    if (lane == 0) v[s] = acc;
  }
  if (blockIdx.x == 0) {
    __shared__ double red[4];
    double q = 0.0;
    for (int64_t p = threadIdx.x; p < np; p += blockDim.x) q += mean[p] * mean[p];
    q = wave_sum(q);
    if (lane == 0) red[wid] = q;
    __syncthreads();
    if (threadIdx.x == 0) c[0] = red[0] + red[1] + red[2] + red[3];
  }
}

// FP32 models: t -> bf16 planes hi + lo (ns x ldt, zero beyond M) for the
// bf16x3 product t t^T, and the spins packed 32 per word (ns x nw).
__global__ void ntk_split_kernel(const float* __restrict__ t, int64_t ns, int M, int ldt,
                                 uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
  const int64_t total = ns * ldt;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e / ldt;
    const int k = (int)(e % ldt);
    uint16_t h = 0, l = 0;
    if (k < M) {
      const float f = t[s * M + k];
      h = bf16_rne(f);
      l = bf16_rne(f - bf16_to_f(h));
    }
    hi[e] = h;
    lo[e] = l;
  }
}

__global__ void ntk_pack_spins_kernel(const int8_t* __restrict__ sigma, int64_t ns, int N, int nw,
                                      uint32_t* __restrict__ bits) {
  const int64_t total = ns * nw;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e / nw;
    const int w = (int)(e % nw);
    const int lim = min(32, N - 32 * w);
    uint32_t word = 0;
    for (int b = 0; b < lim; ++b) word |= (sigma[s * N + 32 * w + b] > 0 ? 1u : 0u) << b;
    bits[e] = word;
  }
}

// Host launchers of the two kernels above, for the other tangent kernels.
void launch_ntk_split(const float* t, int64_t ns, int M, int ldt, uint16_t* hi, uint16_t* lo, hipStream_t s) {
  const unsigned g = (unsigned)std::min<int64_t>((ns * ldt + 255) / 256, 65536);
  hipLaunchKernelGGL(ntk_split_kernel, dim3(g), dim3(256), 0, s, t, ns, M, ldt, hi, lo);
}
void launch_ntk_pack_spins(const int8_t* sigma, int64_t ns, int N, int nw, uint32_t* bits, hipStream_t s) {
  const unsigned g = (unsigned)std::min<int64_t>((ns * nw + 255) / 256, 65536);
  hipLaunchKernelGGL(ntk_pack_spins_kernel, dim3(g), dim3(256), 0, s, sigma, ns, N, nw, bits);
}

// dot = a.b over n entries in one workgroup, summed in a fixed order.
__global__ void __launch_bounds__(1024) dot_f64_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                       int64_t n, double* __restrict__ dot) {
  __shared__ double red[16];
  double q = 0.0;
  for (int64_t p = threadIdx.x; p < n; p += 1024) q += a[p] * b[p];
  q = wave_sum(q);
  if (lane_id() == 0) red[threadIdx.x >> 6] = q;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < 16; ++w) r += red[w];
    dot[0] = r;
  }
}

// y_s = alpha (raw_s - dot) + beta y_s (dot == NULL: 0; beta == 0: y is not read)
__global__ void jvp_finish_kernel(const double* __restrict__ raw, int64_t ns, const double* __restrict__ dot,
                                  double alpha, double beta, double* __restrict__ y) {
  const double d = dot ? dot[0] : 0.0;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += (int64_t)gridDim.x * blockDim.x) {
    const double v = alpha * (raw[s] - d);
    y[s] = beta == 0.0 ? v : v + beta * y[s];
  }
}

// Shared tail of the Jacobian-vector products: y = alpha (raw - mean.v) + beta y.
void launch_jvp_finish(const double* raw, int64_t ns, const double* mean, const double* v, int64_t np,
                       double* dot_slot, double alpha, double beta, double* y, hipStream_t s) {
  if (mean) hipLaunchKernelGGL(dot_f64_kernel, dim3(1), dim3(1024), 0, s, mean, v, np, dot_slot);
  const unsigned g = (unsigned)std::min<int64_t>((ns + 255) / 256, 4096);
  hipLaunchKernelGGL(jvp_finish_kernel, dim3(g), dim3(256), 0, s, raw, ns, mean ? dot_slot : nullptr, alpha,
                     beta, y);
}

bool launch_gram_rbm(const uint16_t* xh, const uint16_t* xl, int64_t n, int64_t K, int64_t ldx,
                     const uint32_t* bits, int nw, int N, const double* v, const double* c, double hb,
                     double ha, double scale, double shift, double offset, double* out, int64_t ldo,
                     hipStream_t s);

void launch_gemm_exact_a(const uint16_t* a, int64_t ma, const uint16_t* bh, const uint16_t* bl, int64_t nb,
                         int64_t K, int64_t ldk, float* c, int64_t ldc, hipStream_t s);

// ---------------------------------------------------------------------------
// Dense S of the RBM from its Khatri-Rao structure (netket/_src/ngd/sr.py:43-55:
// matrix = O_L^T O_L + diag_shift I, O_L = (O - mean) / sqrt(Ns)).
//
// Every parameter's log-derivative is a product O_p = s~_u(p) t~_w(p) of an
// augmented spin s~ = (1, sigma_0..sigma_{N-1}) and an augmented hidden
// activation t~ = (1, tanh theta_0..tanh theta_{M-1}): hidden bias b_j =
// (1, t_j), kernel W_ij = (sigma_i, t_j), visible bias a_i = (sigma_i, 1).  So
//   sum_s O_sp O_sq = sum_s (s~_u s~_u')_s (t~_w t~_w')_s = C[pair(u,u'), pair(w,w')]
// with C = P Q^T over the samples, P the distinct spin-pair products (rows:
// 1, sigma_i, sigma_i sigma_k for i < k: 1 + N + N(N-1)/2 rows, since
// sigma^2 = 1) and Q the distinct activation-pair products (1, t_j, t_j t_l
// for j <= l: 1 + M + M(M+1)/2 rows).  C has about Np^2 / 4 entries against
// the Gram's Np^2 / 2, the spin side is exact in one bf16 plane, and the
// activation side takes two (hi + lo), so the product needs a third of the
// bf16x3 Gram's matrix-core work.  Q is built from the centred activations
// t'_j = t_j - c_j (c_j their sample means, rounded to FP32), so that the
// rounding of the split and of the FP32 accumulation scales with |t'| rather
// than |t|; with c~ = (0, c) the uncentred sum is recovered in FP64 as
//   U(u,u',w,w') = C[r, (w,w')] + c~_w C[r, (-1,w')] + c~_w' C[r, (w,-1)]
//                  + c~_w c~_w' C[r, (-1,-1)],   r = pair(u, u'),
// and centring and the shift are applied when C is unpacked into S:
//   S_pq = scale (U_pq - U_p U_q / ns) (+ shift if p == q),
// U_p = sum_s O_sp = U(u_p, -1, w_p, -1) from the same entries of C.  Taking
// the column sums from the product itself, and not from an exact FP64 mean,
// makes the large terms (c~_w times a pair sum) cancel exactly between U_pq
// and U_p U_q / ns, so the remaining rounding scales with the centred
// activations.  The spin side needs no centring: its pair sums are exact
// integers.
// ---------------------------------------------------------------------------

// Augmented spin index u in [-1, N) (-1: the constant 1).  Row of P holding
// s~_u s~_u'.
__device__ __host__ inline int64_t qgt_spin_row(int u, int v, int N) {
  if (u == v) return 0;
  if (u < 0) return 1 + v;
  if (v < 0) return 1 + u;
  const int64_t i = u < v ? u : v, k = u < v ? v : u;
  return 1 + N + k * (k - 1) / 2 + i;
}
// Augmented hidden index w in [-1, M).  Row of Q holding t~_w t~_w'.
__device__ __host__ inline int64_t qgt_hidden_row(int w, int x, int M) {
  if (w < 0 && x < 0) return 0;
  if (w < 0) return 1 + x;
  if (x < 0) return 1 + w;
  const int64_t j = w < x ? w : x, l = w < x ? x : w;
  return 1 + M + l * (l + 1) / 2 + j;
}
// Largest e with e (e + 1) / 2 <= b.
__device__ inline int64_t tri_root(int64_t b) {
  int64_t e = (int64_t)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((e + 1) * (e + 2) / 2 <= b) ++e;
  while (e * (e + 1) / 2 > b) --e;
  return e;
}
// Position of parameter (u, w) in the flat vector, or -1 when it is not one.
__device__ inline int64_t qgt_param(int u, int w, const nka_rbm& m) {
  if (u < 0) return (w < 0 || m.bias_off < 0) ? -1 : m.bias_off + w;
  if (w < 0) return m.visible_off < 0 ? -1 : m.visible_off + u;
  return m.kernel_off + (int64_t)u * m.n_hidden + w;
}

// sT (N x ldk int8) = sigma^T as +-1, zero for the padding samples s >= ns.
__global__ void qgt_spins_t_kernel(const int8_t* __restrict__ sigma, int64_t ns, int N, int64_t ldk,
                                   int8_t* __restrict__ sT) {
  const int64_t total = (int64_t)N * ldk;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / ldk, s = e % ldk;
    sT[e] = s < ns ? (sigma[s * N + i] > 0 ? (int8_t)1 : (int8_t)-1) : (int8_t)0;
  }
}

__device__ inline uint32_t qgt_sign_bf16(int v) { return v > 0 ? 0x3F80u : (v < 0 ? 0xBF80u : 0u); }

// P (np_rows x ldk bf16): row 0 = 1, row 1 + i = sigma_i, row 1 + N + k(k-1)/2 + i
// = sigma_i sigma_k (i < k); zero for the padding samples.  One workgroup per
// row, eight samples per thread and step.
__global__ void __launch_bounds__(256) qgt_spin_pairs_kernel(const int8_t* __restrict__ sT, int64_t ns, int N,
                                                             int64_t ldk, int64_t rows, uint16_t* __restrict__ P) {
  for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    int64_t i = -1, k = -1;
    if (r >= 1 && r <= N) {
      i = r - 1;
    } else if (r > N) {
      const int64_t b = r - 1 - N;  // k(k-1)/2 + i, i < k
      k = tri_root(b) + 1;
      i = b - k * (k - 1) / 2;
    }
    for (int64_t s0 = (int64_t)threadIdx.x * 8; s0 < ldk; s0 += 256 * 8) {
      int v[8];
      if (i < 0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = s0 + e < ns ? 1 : 0;
      } else {
        const uint2 a = *(const uint2*)(sT + i * ldk + s0);
        const int8_t* ab = (const int8_t*)&a;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = ab[e];
        if (k >= 0) {
          const uint2 b = *(const uint2*)(sT + k * ldk + s0);
          const int8_t* bb = (const int8_t*)&b;
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] *= bb[e];
        }
      }
      uint4 o;
      o.x = qgt_sign_bf16(v[0]) | (qgt_sign_bf16(v[1]) << 16);
      o.y = qgt_sign_bf16(v[2]) | (qgt_sign_bf16(v[3]) << 16);
      o.z = qgt_sign_bf16(v[4]) | (qgt_sign_bf16(v[5]) << 16);
      o.w = qgt_sign_bf16(v[6]) | (qgt_sign_bf16(v[7]) << 16);
      *(uint4*)(P + r * ldk + s0) = o;
    }
  }
}

// c_j = mean over the ns samples of tT[j][:] (FP64, fixed summation order),
// one workgroup per hidden unit.
__global__ void __launch_bounds__(256) qgt_hidden_means_kernel(const float* __restrict__ tT, int64_t ns, int M,
                                                               double* __restrict__ c) {
  __shared__ double red[4];
  for (int j = blockIdx.x; j < M; j += gridDim.x) {
    double acc = 0.0;
    for (int64_t s = threadIdx.x; s < ns; s += 256) acc += (double)tT[(int64_t)j * ns + s];
    acc = wave_sum(acc);
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) c[j] = (red[0] + red[1] + red[2] + red[3]) / (double)ns;
  }
}

// Q hi + lo (rows x ldk bf16 each): row 0 = 1, row 1 + j = t_j, row
// 1 + M + l(l+1)/2 + j = t_j t_l (j <= l, FP32 product), split as
// hi = bf16(q), lo = bf16(q - hi); zero for the padding samples.  tT is
// (M x ns) FP32; t_j here is the centred t_j - (float) c_j.
__global__ void __launch_bounds__(256) qgt_hidden_pairs_kernel(const float* __restrict__ tT, int64_t ns, int M,
                                                               const double* __restrict__ cmean, int64_t ldk,
                                                               int64_t rows, uint16_t* __restrict__ qh,
                                                               uint16_t* __restrict__ ql) {
  const bool vec = (ns % 4) == 0 && (((uintptr_t)tT) % 16) == 0;
  for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
    int64_t j = -1, l = -1;
    if (r >= 1 && r <= M) {
      j = r - 1;
    } else if (r > M) {
      const int64_t b = r - 1 - M;  // l(l+1)/2 + j, j <= l
      l = tri_root(b);
      j = b - l * (l + 1) / 2;
    }
    const float cj = j >= 0 ? (float)cmean[j] : 0.0f, cl = l >= 0 ? (float)cmean[l] : 0.0f;
    for (int64_t s0 = (int64_t)threadIdx.x * 8; s0 < ldk; s0 += 256 * 8) {
      float f[8];
      if (j < 0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) f[e] = s0 + e < ns ? 1.0f : 0.0f;
      } else {
        const float* a = tT + j * ns + s0;
        if (vec && s0 + 8 <= ns) {
          const float4 x = *(const float4*)a, y = *(const float4*)(a + 4);
          f[0] = x.x - cj; f[1] = x.y - cj; f[2] = x.z - cj; f[3] = x.w - cj;
          f[4] = y.x - cj; f[5] = y.y - cj; f[6] = y.z - cj; f[7] = y.w - cj;
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) f[e] = s0 + e < ns ? a[e] - cj : 0.0f;
        }
        if (l >= 0) {
          const float* b = tT + l * ns + s0;
          if (vec && s0 + 8 <= ns) {
            const float4 x = *(const float4*)b, y = *(const float4*)(b + 4);
            f[0] *= x.x - cl; f[1] *= x.y - cl; f[2] *= x.z - cl; f[3] *= x.w - cl;
            f[4] *= y.x - cl; f[5] *= y.y - cl; f[6] *= y.z - cl; f[7] *= y.w - cl;
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) f[e] *= s0 + e < ns ? b[e] - cl : 0.0f;
          }
        }
      }
      uint32_t h[4], lo[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) bf16x2_split(f[2 * e], f[2 * e + 1], h[e], lo[e]);
      const uint4 oh = {h[0], h[1], h[2], h[3]};
      const uint4 ol = {lo[0], lo[1], lo[2], lo[3]};
      *(uint4*)(qh + r * ldk + s0) = oh;
      *(uint4*)(ql + r * ldk + s0) = ol;
    }
  }
}

// S from C, one workgroup per augmented spin pair (u_hi >= u_lo in [-1, N)):
// all entries S[p][q] with p = (u_hi, w), q = (u_lo, w') for w, w' in
// [-1, M), each written once into the lower triangle (at [max][min]).  The
// C row of the pair and the first M + 1 entries of the rows of (u_hi, -1) and
// (u_lo, -1) (the column sums of the two parameter groups) are staged in LDS
// when they fit (LDS_ROW).  Threads run along w', whose positions are
// consecutive in S, so the writes of a row segment are contiguous.
template <bool LDS_ROW>
__global__ void __launch_bounds__(256) qgt_unpack_kernel(nka_rbm m, const float* __restrict__ c, int64_t ldc,
                                                         int64_t nb, const double* __restrict__ cmean, double ns,
                                                         double scale, double shift, double* __restrict__ out,
                                                         int64_t ldo) {
  extern __shared__ float crow[];
  const int N = m.n_visible, M = m.n_hidden;
  const int64_t n_pairs = (int64_t)(N + 1) * (N + 2) / 2;
  const int W = M + 1;
  const double rns = 1.0 / ns;
  const int wid = threadIdx.x >> 6, lane = lane_id();
  for (int64_t b = blockIdx.x; b < n_pairs; b += gridDim.x) {
    const int64_t ah = tri_root(b);
    const int uh = (int)ah - 1, ul = (int)(b - ah * (ah + 1) / 2) - 1;
    const float* src = c + qgt_spin_row(uh, ul, N) * ldc;
    const float* srch = c + qgt_spin_row(uh, -1, N) * ldc;  // entries 0..M: sums of (u_hi, w)
    const float* srcl = c + qgt_spin_row(ul, -1, N) * ldc;
    if (LDS_ROW) {
      __syncthreads();
      for (int64_t e = threadIdx.x; e < nb; e += blockDim.x) crow[e] = src[e];
      for (int e = threadIdx.x; e < W; e += blockDim.x) {
        crow[nb + e] = srch[e];
        crow[nb + W + e] = srcl[e];
      }
      __syncthreads();
    }
    const float* row = LDS_ROW ? crow : src;
    const float* rowh = LDS_ROW ? crow + nb : srch;
    const float* rowl = LDS_ROW ? crow + nb + W : srcl;
    // one element: p = (u_hi, w), q = (u_lo, x); the row's terms are passed in
    auto entry = [&](int w, int64_t p, double cw, double up, int x) {
      const int64_t q = qgt_param(ul, x, m);
      if (q < 0 || (uh == ul && p < q)) return;  // the symmetric diagonal pair: each entry once
      const double cx = x >= 0 ? (double)(float)cmean[x] : 0.0;
      // the uncentred sum from the centred activations' pair sums
      double u = (double)row[qgt_hidden_row(w, x, M)];
      if (w >= 0) u += cw * (double)row[qgt_hidden_row(-1, x, M)];
      if (x >= 0) u += cx * (double)row[qgt_hidden_row(w, -1, M)];
      if (w >= 0 && x >= 0) u += cw * cx * (double)row[0];
      double uq = (double)rowl[x + 1];
      if (x >= 0) uq += cx * (double)rowl[0];
      double v = scale * (u - up * uq * rns);
      if (p == q) v += shift;
      if (p >= q)
        out[p * ldo + q] = v;
      else
        out[q * ldo + p] = v;
    };
    // rows w of the block, one wavefront each; lanes along x >= 0, whose
    // positions q are consecutive in S
    for (int w = wid - 1; w < M; w += 4) {
      const int64_t p = qgt_param(uh, w, m);
      if (p < 0) continue;
      const double cw = w >= 0 ? (double)(float)cmean[w] : 0.0;
      double up = (double)rowh[w + 1];
      if (w >= 0) up += cw * (double)rowh[0];
      for (int x = lane; x < M; x += 64) entry(w, p, cw, up, x);
    }
    // the column x = -1 (the visible bias of u_lo), one row per thread
    for (int w = (int)threadIdx.x - 1; w < M; w += blockDim.x) {
      const int64_t p = qgt_param(uh, w, m);
      if (p < 0) continue;
      const double cw = w >= 0 ? (double)(float)cmean[w] : 0.0;
      double up = (double)rowh[w + 1];
      if (w >= 0) up += cw * (double)rowh[0];
      entry(w, p, cw, up, -1);
    }
  }
}

constexpr int NT_T = 64;
constexpr int NT_K = 16;
constexpr int NT_LD = 17;

// Lower-triangular 64x64 tiles; 4 waves of 32x32 (2x2 v_mfma_f64_16x16x4).
__global__ void __launch_bounds__(256) rbm_ntk_kernel(const double* __restrict__ t64, int M,
                                                      const int8_t* __restrict__ sigma, int N, int64_t ns,
                                                      const double* __restrict__ v,
                                                      const double* __restrict__ c, double hb, double ha,
                                                      double scale, double shift, double offset,
                                                      double* __restrict__ out, int64_t ldo,
                                                      int64_t n_tiles) {
  __shared__ double sA[NT_T * NT_LD];
  __shared__ double sB[NT_T * NT_LD];
  const int64_t G = gridDim.x;
  const int64_t bid = (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;  // XCD-aware order
  if (bid >= n_tiles) return;
  int ti, tj;
  tri_supertile(bid, (ns + NT_T - 1) / NT_T, ti, tj);
  const int64_t i0 = (int64_t)ti * NT_T, j0 = (int64_t)tj * NT_T;
  const int tid = threadIdx.x, wid = tid / 64, lane = lane_id();
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  doublex4 th[2][2], sg[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      th[x][y] = (doublex4){0.0, 0.0, 0.0, 0.0};
      sg[x][y] = (doublex4){0.0, 0.0, 0.0, 0.0};
    }
  // pass 0: Theta = t t^T (K = M); pass 1: Sigma = sigma sigma^T (K = N); the
  // next K slab is loaded into registers while the current one is used
  auto fetch = [&](int pass, int K, int k0, double* pa, double* pb) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + it * 256;
      const int r = idx >> 4, kc = idx & 15;
      const int kk = k0 + kc;
      double va = 0.0, vb = 0.0;
      if (kk < K) {
        if (pass == 0) {
          if (i0 + r < ns) va = t64[(i0 + r) * M + kk];
          if (j0 + r < ns) vb = t64[(j0 + r) * M + kk];
        } else {
          if (i0 + r < ns) va = sigma[(i0 + r) * N + kk] > 0 ? 1.0 : -1.0;
          if (j0 + r < ns) vb = sigma[(j0 + r) * N + kk] > 0 ? 1.0 : -1.0;
        }
      }
      pa[it] = va;
      pb[it] = vb;
    }
  };
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const int K = pass == 0 ? M : N;
    double pa[4], pb[4];
    fetch(pass, K, 0, pa, pb);
    for (int k0 = 0; k0 < K; k0 += NT_K) {
      __syncthreads();
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + it * 256;
        const int r = idx >> 4, kc = idx & 15;
        sA[r * NT_LD + kc] = pa[it];
        sB[r * NT_LD + kc] = pb[it];
      }
      __syncthreads();
      if (k0 + NT_K < K) fetch(pass, K, k0 + NT_K, pa, pb);
#pragma unroll
      for (int ks = 0; ks < NT_K; ks += 4) {
        const int kq = ks + (lane >> 4);
        double av[2], bv[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) av[x] = sA[(wm + x * 16 + (lane & 15)) * NT_LD + kq];
#pragma unroll
        for (int y = 0; y < 2; ++y) bv[y] = sB[(wn + y * 16 + (lane & 15)) * NT_LD + kq];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            if (pass == 0)
              th[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], th[x][y], 0, 0, 0);
            else
              sg[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], sg[x][y], 0, 0, 0);
          }
      }
    }
  }
  const double cc = *c;
  // element e of lane l in a 16x16 tile is (row l/16 + 4e, col l%16)
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t gi = i0 + wm + x * 16 + (lane >> 4) + 4 * e;
        const int64_t gj = j0 + wn + y * 16 + (lane & 15);
        if (gi < ns && gj <= gi) {
          const double S = sg[x][y][e], Th = th[x][y][e];
          double val = scale * ((hb + S) * Th + ha * S - v[gi] - v[gj] + cc) + offset;
          if (gi == gj) val += shift;
          out[gi * ldo + gj] = val;
        }
      }
}

}  // namespace nka

using namespace nka;

extern "C" {

static int64_t ntk_base_bytes(const nka_rbm& m, int64_t ns) {
  return (ns * (int64_t)m.n_hidden + ns * (int64_t)m.n_visible + ns + 8) * (int64_t)sizeof(double);
}
static int ntk_ldt(const nka_rbm& m) { return (m.n_hidden + 7) / 8 * 8; }
static int ntk_nw(const nka_rbm& m) { return (m.n_visible + 31) / 32; }

int64_t nka_rbm_ntk_workspace(const nka_rbm* m, int64_t ns) {
  if (!m || ns <= 0) return 0;
  int64_t b = ntk_base_bytes(*m, ns);
  if (m->dtype == NKA_F32) {
    b = (b + 255) / 256 * 256;
    b += (2 * ns * ntk_ldt(*m) * (int64_t)sizeof(uint16_t) + 255) / 256 * 256;
    b += ns * ntk_nw(*m) * (int64_t)sizeof(uint32_t);
  }
  return b;
}

int64_t nka_rbm_jvp_workspace(const nka_rbm* m, int64_t ns) {
  if (!m || ns <= 0) return 0;
  return ntk_base_bytes(*m, ns) + 256;
}

// (O v)_s = t_s.v_b + sum_i sigma_si (t_s v_W^T)_i + sigma_s.v_a: the product
// t v_W^T on FP64 matrix cores and the per-sample sums of the NTK's v = O mu
// with v in place of mu.
int nka_rbm_jvp(const nka_rbm* m, const int8_t* sigma, const void* t, int64_t ns, const double* v,
                const double* mean, double alpha, double beta, double* y, void* workspace, void* stream) {
  NKA_CHECK_ARG(m && m->params && sigma && t && v && y && workspace && ns > 0, "nka_rbm_jvp: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int N = m->n_visible, M = m->n_hidden;
  const int64_t np = (m->bias_off >= 0 ? M : 0) + (int64_t)N * M + (m->visible_off >= 0 ? N : 0);
  double* t64 = (double*)workspace;
  double* U = t64 + ns * M;
  double* raw = U + ns * N;
  double* c = raw + ns;  // |v|^2 (written by the shared kernel, not used) and mean.v
  const int64_t total = ns * M;
  const unsigned g1 = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
  if (m->dtype == NKA_F32)
    hipLaunchKernelGGL(to_f64_kernel<float>, dim3(g1), dim3(256), 0, s, (const float*)t, total, t64);
  else
    hipLaunchKernelGGL(to_f64_kernel<double>, dim3(g1), dim3(256), 0, s, (const double*)t, total, t64);
  int st = nka_gemm_nt_f64(t64, M, v + m->kernel_off, M, U, N, ns, N, M, 1.0, 0.0, stream);
  if (st) return st;
  const unsigned g2 = (unsigned)std::min<int64_t>((ns + 3) / 4, 4096);
  hipLaunchKernelGGL(ntk_vec_kernel, dim3(g2), dim3(256), 0, s, t64, sigma, U, ns, N, M, v, m->bias_off,
                     m->visible_off, np, raw, c);
  launch_jvp_finish(raw, ns, mean, v, np, c + 1, alpha, beta, y, s);
  NKA_CHECK_LAUNCH("nka_rbm_jvp");
  return NKA_OK;
}

int nka_rbm_ntk(const nka_rbm* m, const int8_t* sigma, const void* t, int64_t ns,
                const double* mean, double scale, double shift, double offset, double* out,
                int64_t ldo, void* workspace, void* stream) {
  NKA_CHECK_ARG(m && m->params && sigma && t && mean && out && workspace && ns > 0,
                "nka_rbm_ntk: bad arguments");
  NKA_CHECK_ARG(ldo >= ns, "nka_rbm_ntk: ldo < ns");
  hipStream_t s = (hipStream_t)stream;
  const int N = m->n_visible, M = m->n_hidden;
  const int64_t np = (m->bias_off >= 0 ? M : 0) + (int64_t)N * M + (m->visible_off >= 0 ? N : 0);
  double* t64 = (double*)workspace;
  double* U = t64 + ns * M;
  double* v = U + ns * N;
  double* c = v + ns;
  const int64_t total = ns * M;
  const unsigned g1 = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
  if (m->dtype == NKA_F32)
    hipLaunchKernelGGL(to_f64_kernel<float>, dim3(g1), dim3(256), 0, s, (const float*)t, total, t64);
  else
    hipLaunchKernelGGL(to_f64_kernel<double>, dim3(g1), dim3(256), 0, s, (const double*)t, total, t64);
  // U = t mu_W^T (ns x N): mu_W is the (N x M) block of the means
  int st = nka_gemm_nt_f64(t64, M, mean + m->kernel_off, M, U, N, ns, N, M, 1.0, 0.0, stream);
  if (st) return st;
  const unsigned g2 = (unsigned)std::min<int64_t>((ns + 3) / 4, 4096);
  hipLaunchKernelGGL(ntk_vec_kernel, dim3(g2), dim3(256), 0, s, t64, sigma, U, ns, N, M, mean,
                     m->bias_off, m->visible_off, np, v, c);
  const double hb = m->bias_off >= 0 ? 1.0 : 0.0, ha = m->visible_off >= 0 ? 1.0 : 0.0;
  if (m->dtype == NKA_F32) {
    // Theta = t t^T on bf16 matrix cores (hi hi^T + hi lo^T + lo hi^T, FP32
    // accumulation); S and the centring are exact in the epilogue
    unsigned char* p = (unsigned char*)workspace + (ntk_base_bytes(*m, ns) + 255) / 256 * 256;
    const int ldt = ntk_ldt(*m), nw = ntk_nw(*m);
    uint16_t* hi = (uint16_t*)p;
    uint16_t* lo = hi + ns * ldt;
    uint32_t* bits = (uint32_t*)(p + (2 * ns * ldt * (int64_t)sizeof(uint16_t) + 255) / 256 * 256);
    const unsigned gs = (unsigned)std::min<int64_t>((ns * ldt + 255) / 256, 65536);
    hipLaunchKernelGGL(ntk_split_kernel, dim3(gs), dim3(256), 0, s, (const float*)t, ns, M, ldt, hi, lo);
    const unsigned gb = (unsigned)std::min<int64_t>((ns * nw + 255) / 256, 65536);
    hipLaunchKernelGGL(ntk_pack_spins_kernel, dim3(gb), dim3(256), 0, s, sigma, ns, N, nw, bits);
    const int64_t T2 = (ns + 127) / 128;
    NKA_CHECK_ARG(T2 * (T2 + 1) / 2 < (1LL << 31), "nka_rbm_ntk: too large");
    if (launch_gram_rbm(hi, lo, ns, ldt, ldt, bits, nw, N, v, c, hb, ha, scale, shift, offset, out, ldo, s)) {
      NKA_CHECK_LAUNCH("nka_rbm_ntk");
      return NKA_OK;
    }
  }
  const int64_t T = (ns + NT_T - 1) / NT_T;
  const int64_t tiles = T * (T + 1) / 2;
  NKA_CHECK_ARG(tiles < (1LL << 31), "nka_rbm_ntk: too large");
  const unsigned grid = (unsigned)(((tiles + 7) / 8) * 8);
  hipLaunchKernelGGL(rbm_ntk_kernel, dim3(grid), dim3(256), 0, s, t64, M, sigma, N, ns, v, c, hb, ha,
                     scale, shift, offset, out, ldo, tiles);
  NKA_CHECK_LAUNCH("nka_rbm_ntk");
  return NKA_OK;
}

// Workspace of nka_rbm_qgt: transposed spins, P, Q hi and lo, and C.
static int64_t qgt_ldk(int64_t ns) { return (ns + 63) / 64 * 64; }
static int64_t qgt_rows_p(const nka_rbm& m) { return 1 + (int64_t)m.n_visible + (int64_t)m.n_visible * (m.n_visible - 1) / 2; }
static int64_t qgt_rows_q(const nka_rbm& m) { return 1 + (int64_t)m.n_hidden + (int64_t)m.n_hidden * (m.n_hidden + 1) / 2; }
static int64_t qgt_ldc(const nka_rbm& m) { return (qgt_rows_q(m) + 63) / 64 * 64; }
static int64_t qgt_round(int64_t b) { return (b + 255) / 256 * 256; }

int64_t nka_rbm_qgt_workspace(const nka_rbm* m, int64_t ns) {
  if (!m || ns <= 0) return 0;
  const int64_t ldk = qgt_ldk(ns);
  return qgt_round((int64_t)m->n_visible * ldk) + qgt_round(qgt_rows_p(*m) * ldk * 2) +
         2 * qgt_round(qgt_rows_q(*m) * ldk * 2) + qgt_round(qgt_rows_p(*m) * qgt_ldc(*m) * 4) +
         qgt_round((int64_t)m->n_hidden * 8);
}

int nka_rbm_qgt(const nka_rbm* m, const int8_t* sigma, const void* tT, int64_t ns, double scale, double shift,
                double* out, int64_t ldo, void* workspace, void* stream) {
  NKA_CHECK_ARG(m && m->params && sigma && tT && out && workspace && ns > 0,
                "nka_rbm_qgt: bad arguments");
  NKA_CHECK_ARG(m->dtype == NKA_F32, "nka_rbm_qgt: FP32 models only (FP64 models take the FP64 Jacobian)");
  const int N = m->n_visible, M = m->n_hidden;
  const int64_t np = (m->bias_off >= 0 ? M : 0) + (int64_t)N * M + (m->visible_off >= 0 ? N : 0);
  NKA_CHECK_ARG(ldo >= np, "nka_rbm_qgt: ldo < n_params");
  NKA_CHECK_ARG((int64_t)(N + 1) * (N + 2) / 2 < (1LL << 31), "nka_rbm_qgt: too many visible units");
  hipStream_t s = (hipStream_t)stream;
  const int64_t ldk = qgt_ldk(ns), rp = qgt_rows_p(*m), rq = qgt_rows_q(*m), ldc = qgt_ldc(*m);
  unsigned char* w = (unsigned char*)workspace;
  int8_t* sT = (int8_t*)w;
  w += qgt_round((int64_t)N * ldk);
  uint16_t* P = (uint16_t*)w;
  w += qgt_round(rp * ldk * 2);
  uint16_t* qh = (uint16_t*)w;
  w += qgt_round(rq * ldk * 2);
  uint16_t* ql = (uint16_t*)w;
  w += qgt_round(rq * ldk * 2);
  float* C = (float*)w;
  w += qgt_round(rp * ldc * 4);
  double* cmean = (double*)w;
  const unsigned g0 = (unsigned)std::min<int64_t>(((int64_t)N * ldk + 255) / 256, 65536);
  hipLaunchKernelGGL(qgt_spins_t_kernel, dim3(g0), dim3(256), 0, s, sigma, ns, N, ldk, sT);
  hipLaunchKernelGGL(qgt_spin_pairs_kernel, dim3((unsigned)std::min<int64_t>(rp, 65536)), dim3(256), 0, s, sT,
                     ns, N, ldk, rp, P);
  hipLaunchKernelGGL(qgt_hidden_means_kernel, dim3((unsigned)std::min<int64_t>(M, 65536)), dim3(256), 0, s,
                     (const float*)tT, ns, M, cmean);
  hipLaunchKernelGGL(qgt_hidden_pairs_kernel, dim3((unsigned)std::min<int64_t>(rq, 65536)), dim3(256), 0, s,
                     (const float*)tT, ns, M, (const double*)cmean, ldk, rq, qh, ql);
  launch_gemm_exact_a(P, rp, qh, ql, rq, ldk, ldk, C, ldc, s);
  const int64_t pairs = (int64_t)(N + 1) * (N + 2) / 2;
  const unsigned gu = (unsigned)std::min<int64_t>(pairs, 65536);
  const size_t row_bytes = (size_t)(rq + 2 * (M + 1)) * sizeof(float);
  if (row_bytes <= 64 * 1024)
    hipLaunchKernelGGL(qgt_unpack_kernel<true>, dim3(gu), dim3(256), row_bytes, s, *m, C, ldc, rq,
                       (const double*)cmean, (double)ns, scale, shift, out, ldo);
  else
    hipLaunchKernelGGL(qgt_unpack_kernel<false>, dim3(gu), dim3(256), 0, s, *m, C, ldc, rq, (const double*)cmean,
                       (double)ns, scale, shift, out, ldo);
  NKA_CHECK_LAUNCH("nka_rbm_qgt");
  return NKA_OK;
}

}  // extern "C"
