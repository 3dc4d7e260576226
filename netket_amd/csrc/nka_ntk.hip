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

}  // extern "C"
