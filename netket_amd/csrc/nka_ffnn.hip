// Two-layer dense network ansatz on gfx950:
//   log psi(x) = sum_k log_cosh(z2_k),  z2 = log_cosh(x W1 + b1) W2 + b2
// (Dense -> log_cosh -> Dense -> log_cosh -> sum; nn.Dense of
// netket/nn/blocks/mlp.py:105-115, log_cosh of netket/nn/activation.py:78-84).
//
// Every evaluation works on tiles of 32 configurations per workgroup: the
// hidden layer h1 of the tile sits in LDS as the A operand of the second
// layer, whose W2 operand streams from L2 (each element read once per
// tile).  FP32 models use v_mfma_f32_32x32x2f32 for that product; FP64
// models use plain FMAs.  Row sums are reduced in a fixed order (butterfly
// shuffles, then per-block partials summed in order), so results do not
// depend on scheduling.
//
// Callers of the tile kernel: log psi, the local energies (connected
// configurations from the conn kernels, first layer updated from the
// sample's theta1 by the differing sites), the per-sample quantities of the
// Jacobian (h1, tanh z2, back-propagated d1) and the Metropolis sampler
// (32 chains per workgroup stepping in lock step).
#include "nka_common.hpp"
#include "netket_amd.h"

extern "C" int nka_ising_conn(const int8_t* x, int64_t n_batch, int n_sites, const int32_t* edges,
                              int n_edges, double h, double J, int mel_dtype, int n_rows, int8_t* xp,
                              void* mels, void* stream);
extern "C" int nka_localop_conn(const int8_t* x, int64_t n_batch, int n_sites, const nka_localop* op,
                                int8_t* xp, void* mels, int32_t* n_conn, void* stream);
extern "C" int nka_gemm_nt_f64(const double* a, int64_t lda, const double* b, int64_t ldb, double* c,
                               int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta,
                               void* stream);

namespace nka {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int FF_ROWS = 32;      // configurations (or chains) per workgroup
constexpr int FF_THREADS = 256;
constexpr int FF_GROUPS = FF_THREADS / FF_ROWS;  // threads per row in the first layer
constexpr int FF_HPT = 128;      // hidden units per thread: H1 <= FF_GROUPS * FF_HPT
constexpr int FF_MAXH = FF_GROUPS * FF_HPT;      // 1024
constexpr int FF_MAXBLK = FF_MAXH / 32;          // column blocks of the second layer
constexpr int FF_MAXD = 4;                        // differing sites per connected row, fast path

template <class T>
struct FfnnView {
  int N, H1, H2;
  const T* b1;
  const T* W1;
  const T* b2;
  const T* W2;
};

template <class T>
__host__ __device__ inline FfnnView<T> ffnn_view(const nka_ffnn& m) {
  const T* p = (const T*)m.params;
  FfnnView<T> v;
  v.N = m.n_visible;
  v.H1 = m.n_hidden1;
  v.H2 = m.n_hidden2;
  v.b1 = m.b1_off >= 0 ? p + m.b1_off : nullptr;
  v.W1 = p + m.w1_off;
  v.b2 = m.b2_off >= 0 ? p + m.b2_off : nullptr;
  v.W2 = p + m.w2_off;
  return v;
}

__host__ __device__ inline int lda_of(int H1) { return H1 + ((H1 & 1) ? 2 : 1); }  // odd, >= H1 + 1
__host__ __device__ inline int64_t ffpad(int64_t b) { return (b + 15) & ~(int64_t)15; }
// LDS scratch of rows_gemm's K-split blocks for ncols output columns (FP32)
__host__ __device__ inline int64_t split_scratch_bytes(int ncols) { return (int64_t)(((ncols + 31) / 32) & 3) * 4 * 1024 * 4; }
constexpr int64_t FF_LDS_MAX = 160 * 1024;

// ---------------------------------------------------------------------------
// Second layer over a 32-row tile: z[r][c] = sum_k A[r][k] B(k, c) for the
// rows of A (LDS, row stride lda, K columns) and ncols columns; epi(r, c, z)
// is applied to every element and its values are summed per row into
// rowsum[r] (0 <= r < 32) in a fixed order.  part is LDS scratch of
// 32 * FF_MAXBLK values.  All threads of the workgroup must call it.
// The 32-column blocks go to the four wavefronts in turn; with rbuf (LDS
// scratch of (nblk % 4) * 4 * 1024 values) the nblk % 4 blocks left over
// after whole rounds are split along K over the four wavefronts instead, and
// their partial tiles are summed in a fixed order, so every wavefront does the
// same number of products.
// ---------------------------------------------------------------------------
template <class Epi>
__device__ void rows_gemm(const float* As, int lda, int K, int ncols, const float* __restrict__ Bp, int sk,
                          int sc, Epi epi, float* part, float* rowsum, float* rbuf = nullptr) {
  const int wid = threadIdx.x >> 6, lane = lane_id();
  const int nblk = (ncols + 31) >> 5;
  const int kh = lane >> 5, lc = lane & 31;
  constexpr int KB = 16;
  // acc += A[:, ka:kb] B(ka:kb, block cb); lane (lc, kh) supplies row lc and
  // column cb*32 + lc at k = k0 + 2q + kh of each pair of k-rows.  The B
  // loads run three groups of 16 k-rows ahead of their products (three
  // register sets in turn; near the end of the range a group reloads the
  // last full group, unused, so that every load is unconditional).
  auto mfma_range = [&](floatx16& acc, int cb, int ka, int kb) {
    const int col = cb * 32 + lc;
    const bool cok = col < ncols;
    const int colc = cok ? col : ncols - 1;
    const uint32_t boff = (uint32_t)colc * (uint32_t)sc + (uint32_t)kh * (uint32_t)sk;
    const uint32_t s2 = 2u * (uint32_t)sk;
    const int kend = ka + (kb - ka) / KB * KB;
    int k = ka;
    auto loadB = [&](int kk0, float* bv) {
      const int kc = kk0 < kend ? kk0 : kend - KB;
      const uint32_t kn = (uint32_t)kc * (uint32_t)sk + boff;
#pragma unroll
      for (int q = 0; q < KB / 2; ++q) bv[q] = Bp[kn + (uint32_t)q * s2];
    };
    auto group = [&](int kk0, const float* bv) {
      float a[KB / 2];
#pragma unroll
      for (int q = 0; q < KB / 2; ++q) a[q] = As[lc * lda + kk0 + 2 * q + kh];
#pragma unroll
      for (int q = 0; q < KB / 2; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q], bv[q], acc, 0, 0, 0);
    };
    if (kend > ka) {
      float b0[KB / 2], b1[KB / 2], b2[KB / 2];
      loadB(k, b0);
      loadB(k + KB, b1);
      while (true) {
        loadB(k + 2 * KB, b2);
        __builtin_amdgcn_sched_barrier(0);
        group(k, b0);
        k += KB;
        if (k >= kend) break;
        loadB(k + 2 * KB, b0);
        __builtin_amdgcn_sched_barrier(0);
        group(k, b1);
        k += KB;
        if (k >= kend) break;
        loadB(k + 2 * KB, b1);
        __builtin_amdgcn_sched_barrier(0);
        group(k, b2);
        k += KB;
        if (k >= kend) break;
      }
    }
    for (; k < kb; k += 2) {
      const int kk = k + kh;
      const float av = kk < kb ? As[lc * lda + kk] : 0.0f;
      const float bv = (kk < kb && cok) ? Bp[(uint32_t)kk * (uint32_t)sk + (uint32_t)col * (uint32_t)sc] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
    }
  };
  // element e of lane l is (row (e/4)*8 + (l/32)*4 + e%4, col l%32)
  auto finish = [&](const floatx16& acc, int cb) {
    const int col = cb * 32 + lc;
    const bool cok = col < ncols;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int row = (e >> 2) * 8 + kh * 4 + (e & 3);
      float v = cok ? epi(row, col, acc[e]) : 0.0f;
#pragma unroll
      for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lc == 0) part[row * FF_MAXBLK + cb] = v;
    }
  };
  const int rem = rbuf ? (nblk & 3) : 0;
  const int nfull = nblk - rem;
  for (int cb = wid; cb < nfull; cb += 4) {
    floatx16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    mfma_range(acc, cb, 0, K);
    finish(acc, cb);
  }
  if (rem > 0) {
    // K quarters rounded up to whole groups; wavefront w takes quarter w
    const int kq = ((K + 3) / 4 + KB - 1) / KB * KB;
    const int k0 = min(K, wid * kq), k1 = min(K, k0 + kq);
    for (int j = 0; j < rem; ++j) {
      floatx16 acc;
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
      mfma_range(acc, nfull + j, k0, k1);
      float* dst = rbuf + (j * 4 + wid) * 1024;
#pragma unroll
      for (int e = 0; e < 16; ++e) dst[e * 64 + lane] = acc[e];
    }
    __syncthreads();
    if (wid < rem) {
      floatx16 acc;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        float t = 0.0f;
#pragma unroll
        for (int w = 0; w < 4; ++w) t += rbuf[(wid * 4 + w) * 1024 + e * 64 + lane];
        acc[e] = t;
      }
      finish(acc, nfull + wid);
    }
  }
  __syncthreads();
  if (threadIdx.x < FF_ROWS) {
    float s = 0.0f;
    for (int cb = 0; cb < nblk; ++cb) s += part[threadIdx.x * FF_MAXBLK + cb];
    rowsum[threadIdx.x] = s;
  }
  __syncthreads();
}

template <class Epi>
__device__ void rows_gemm(const double* As, int lda, int K, int ncols, const double* __restrict__ Bp, int sk,
                          int sc, Epi epi, double* part, double* rowsum, double* rbuf = nullptr) {
  const int r = threadIdx.x & 31, g = threadIdx.x >> 5;
  double s = 0.0;
  for (int c = g; c < ncols; c += FF_GROUPS) {
    double z = 0.0;
    for (int k = 0; k < K; ++k) z += As[r * lda + k] * Bp[(int64_t)k * sk + (int64_t)c * sc];
    s += epi(r, c, z);
  }
  part[r * FF_MAXBLK + g] = s;
  __syncthreads();
  if (threadIdx.x < FF_ROWS) {
    double t = 0.0;
    for (int q = 0; q < FF_GROUPS; ++q) t += part[threadIdx.x * FF_MAXBLK + q];
    rowsum[threadIdx.x] = t;
  }
  __syncthreads();
}

// theta1 of unit a for a configuration given as int8 spins in LDS
template <class T>
__device__ inline T theta1_full(const FfnnView<T>& v, const int8_t* x, int a) {
  T t = v.b1 ? v.b1[a] : (T)0;
  for (int i = 0; i < v.N; ++i) t += (x[i] > 0 ? (T)1 : (T)-1) * v.W1[(int64_t)i * v.H1 + a];
  return t;
}

// ---------------------------------------------------------------------------
// Batched forward pass.  MODE 0: rows are configurations x[r]; theta1 from
// scratch (optionally stored to th_out).  MODE 1: rows r = (s, k) of
// connected configurations xp[r] of sample s = r / K, taken from the list
// ridx[0 .. *rcount) (the rows with a nonzero matrix element); theta1 is the
// sample's theta1 (th_base[s]) plus 2 x'_i W1[i] for every site where x'
// differs from x[s].  Outputs log psi per row (MODE 1: at the row's own
// index); g2_out (rows x H2), when set, receives tanh(z2).
// ---------------------------------------------------------------------------
template <class T, int MODE>
__global__ void __launch_bounds__(FF_THREADS) ffnn_forward_kernel(nka_ffnn m, const int8_t* __restrict__ xr,
                                                                 const int8_t* __restrict__ xs, int Kc,
                                                                 const T* __restrict__ th_base, int64_t rows,
                                                                 T* __restrict__ logpsi, T* __restrict__ th_out,
                                                                 T* __restrict__ g2_out, const int* __restrict__ ridx,
                                                                 const int* __restrict__ rcount) {
  extern __shared__ unsigned char smem[];
  const FfnnView<T> v = ffnn_view<T>(m);
  const int N = v.N, H1 = v.H1, H2 = v.H2;
  const int lda = lda_of(H1);
  T* hA = (T*)smem;                                   // 32 x lda
  T* part = hA + FF_ROWS * lda;                        // 32 x FF_MAXBLK
  T* rowsum = part + FF_ROWS * FF_MAXBLK;              // 32
  int* dsite = (int*)(rowsum + FF_ROWS);               // 32 x (FF_MAXD + 1): count, sites
  int8_t* xl = (int8_t*)(dsite + FF_ROWS * (FF_MAXD + 1));  // 32 x N (+ 32 x N base rows)
  int8_t* xb = xl + FF_ROWS * N;
  T* rbuf = sizeof(T) == 4 ? (T*)(smem + ffpad((int64_t)(xb - (int8_t*)smem) + (int64_t)FF_ROWS * N)) : nullptr;
  const int64_t r0 = (int64_t)blockIdx.x * FF_ROWS;
  // MODE 1: tile rows r0 .. r0+31 of the row list; the grid covers the
  // longest possible list, so tiles past its end leave at once
  const int64_t nrows = MODE == 1 ? (int64_t)*rcount : rows;
  if (r0 >= nrows) return;
  auto row_of = [&](int r) -> int64_t {
    const int64_t cr = r0 + r;
    if (cr >= nrows) return -1;
    return MODE == 1 ? (int64_t)ridx[cr] : cr;
  };
  for (int idx = threadIdx.x; idx < FF_ROWS * N; idx += FF_THREADS) {
    const int r = idx / N, i = idx % N;
    const int64_t gr = row_of(r);
    xl[idx] = gr >= 0 ? xr[gr * N + i] : (int8_t)1;
    if (MODE == 1) xb[idx] = gr >= 0 ? xs[(gr / Kc) * N + i] : (int8_t)1;
  }
  __syncthreads();
  if (MODE == 1) {
    // the sites where row r differs from its sample, ascending; a count of
    // -1 marks a row with more than FF_MAXD of them (full scan below)
    if (threadIdx.x < FF_ROWS) {
      const int rr = threadIdx.x;
      int cnt = 0;
      for (int i = 0; i < N; ++i)
        if (xl[rr * N + i] != xb[rr * N + i]) {
          if (cnt < FF_MAXD) dsite[rr * (FF_MAXD + 1) + 1 + cnt] = i;
          ++cnt;
        }
      dsite[rr * (FF_MAXD + 1)] = cnt <= FF_MAXD ? cnt : -1;
    }
    __syncthreads();
  }
  const int r = threadIdx.x / FF_GROUPS, g = threadIdx.x % FF_GROUPS;
  const int64_t gr = row_of(r);
  int nd = 0, ds[FF_MAXD];
  T sg[FF_MAXD];
  if (MODE == 1) {
    nd = dsite[r * (FF_MAXD + 1)];
#pragma unroll
    for (int d = 0; d < FF_MAXD; ++d) {
      ds[d] = d < nd ? dsite[r * (FF_MAXD + 1) + 1 + d] : 0;
      sg[d] = xl[r * N + ds[d]] > 0 ? (T)2 : (T)-2;
    }
  }
  for (int a = g; a < H1; a += FF_GROUPS) {
    T t;
    if (MODE == 0) {
      t = theta1_full<T>(v, xl + r * N, a);
      if (th_out && gr >= 0) th_out[gr * H1 + a] = t;
    } else {
      t = gr >= 0 ? th_base[(gr / Kc) * H1 + a] : (T)0;
      if (nd >= 0) {
        // same terms in the same (ascending) order as the full scan
#pragma unroll
        for (int d = 0; d < FF_MAXD; ++d)
          if (d < nd) t += sg[d] * v.W1[(int64_t)ds[d] * H1 + a];
      } else {
        const int8_t* x1 = xl + r * N;
        const int8_t* x0 = xb + r * N;
        for (int i = 0; i < N; ++i)
          if (x1[i] != x0[i]) t += (x1[i] > 0 ? (T)2 : (T)-2) * v.W1[(int64_t)i * H1 + a];
      }
    }
    hA[r * lda + a] = log_cosh<T>(t);
  }
  __syncthreads();
  const T* W2 = v.W2;
  const T* b2 = v.b2;
  auto epi = [=](int rr, int c, T z) {
    const T zz = z + (b2 ? b2[c] : (T)0);
    if (g2_out && r0 + rr < rows) g2_out[(r0 + rr) * H2 + c] = tanh_t<T>(zz);
    return log_cosh<T>(zz);
  };
  rows_gemm(hA, lda, H1, H2, W2, H2, 1, epi, part, rowsum, rbuf);
  if (threadIdx.x < FF_ROWS) {
    const int64_t go = row_of(threadIdx.x);
    if (go >= 0) logpsi[go] = rowsum[threadIdx.x];
  }
}

template <class T>
inline size_t forward_lds(const nka_ffnn& m) {
  return ffpad((size_t)FF_ROWS * lda_of(m.n_hidden1) * sizeof(T)) +
         ffpad((size_t)FF_ROWS * FF_MAXBLK * sizeof(T)) + ffpad(FF_ROWS * sizeof(T)) +
         ffpad((size_t)FF_ROWS * (FF_MAXD + 1) * sizeof(int)) + ffpad((size_t)2 * FF_ROWS * m.n_visible) +
         (sizeof(T) == 4 ? split_scratch_bytes(m.n_hidden2) : 0);
}

// List of the connected rows with a nonzero matrix element (only those enter
// the local energy; the reduction below skips the others).  Slots are taken by
// one atomic add per wavefront, so the list order varies between runs, but
// every row's log psi is formed from that row alone, so the results do not.
__global__ void ffnn_zero_count_kernel(int* __restrict__ count) {
  if (threadIdx.x == 0) count[0] = 0;
}

template <class M>
__global__ void __launch_bounds__(256) ffnn_compact_rows_kernel(const M* __restrict__ mels, int64_t rows,
                                                                int* __restrict__ ridx, int* __restrict__ count) {
  const int lane = lane_id();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < rows; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = base + threadIdx.x;
    const bool nz = r < rows && mels[r] != (M)0;
    const uint64_t bal = __ballot(nz);
    int slot0 = 0;
    if (lane == 0 && bal) slot0 = atomicAdd(count, (int)__popcll(bal));
    slot0 = __shfl(slot0, 0, 64);
    if (nz) ridx[slot0 + (int)__popcll(bal & ((1ull << lane) - 1ull))] = (int)r;
  }
}

// eloc[s] = sum_k mel[s,k] exp(lp[s,k] - lp0[s]) (FP64); one wavefront per sample
template <class T, class M>
__global__ void ffnn_eloc_reduce_kernel(const M* __restrict__ mels, const T* __restrict__ lp,
                                        const T* __restrict__ lp0, int64_t ns, int K, double* __restrict__ eloc) {
  const int wid = threadIdx.x >> 6, lane = lane_id();
  const int64_t s = (int64_t)blockIdx.x * (blockDim.x >> 6) + wid;
  if (s >= ns) return;
  const double l0 = (double)lp0[s];
  double acc = 0.0;
  for (int k = lane; k < K; k += 64) {
    const double mel = (double)mels[s * K + k];
    if (mel != 0.0) acc += mel * exp((double)lp[s * K + k] - l0);
  }
  acc = wave_sum(acc);
  if (lane == 0) eloc[s] = acc;
}

// ---------------------------------------------------------------------------
// Per-sample quantities of the log-derivatives: h1 = log_cosh(theta1),
// g2 = tanh(z2) (from the forward pass) and d1 = (W2 g2) * tanh(theta1), the
// gradient with respect to theta1 (a second tile product with B = W2^T).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(FF_THREADS) ffnn_backprop_kernel(nka_ffnn m, const T* __restrict__ th1,
                                                                  const T* __restrict__ g2, int64_t rows,
                                                                  T* __restrict__ h1_out, T* __restrict__ d1_out) {
  extern __shared__ unsigned char smem[];
  const FfnnView<T> v = ffnn_view<T>(m);
  const int H1 = v.H1, H2 = v.H2;
  const int ldg = lda_of(H2);
  T* gA = (T*)smem;                           // 32 x ldg
  T* part = gA + FF_ROWS * ldg;
  T* rowsum = part + FF_ROWS * FF_MAXBLK;
  const int64_t r0 = (int64_t)blockIdx.x * FF_ROWS;
  for (int idx = threadIdx.x; idx < FF_ROWS * H2; idx += FF_THREADS) {
    const int r = idx / H2, c = idx % H2;
    gA[r * ldg + c] = r0 + r < rows ? g2[(r0 + r) * H2 + c] : (T)0;
  }
  for (int idx = threadIdx.x; idx < FF_ROWS * H1; idx += FF_THREADS) {
    const int r = idx / H1, a = idx % H1;
    if (r0 + r < rows) h1_out[(r0 + r) * H1 + a] = log_cosh<T>(th1[(r0 + r) * H1 + a]);
  }
  __syncthreads();
  const T* W2 = v.W2;
  auto epi = [=](int rr, int a, T u) {
    if (r0 + rr < rows) d1_out[(r0 + rr) * H1 + a] = u * tanh_t<T>(th1[(r0 + rr) * H1 + a]);
    return (T)0;
  };
  rows_gemm(gA, ldg, H2, H1, W2, 1, H2, epi, part, rowsum);
}

template <class T>
inline size_t backprop_lds(const nka_ffnn& m) {
  return ffpad((size_t)FF_ROWS * lda_of(m.n_hidden2) * sizeof(T)) +
         ffpad((size_t)FF_ROWS * FF_MAXBLK * sizeof(T)) + ffpad(FF_ROWS * sizeof(T));
}

// ---------------------------------------------------------------------------
// Weighted sums over samples of the log-derivatives (FP64):
//   out[v, p] = sum_s (O[s, p] - mean[p]) w[v, s]
// with O = [d1 | x (x) d1 | g2 | h1 (x) g2] in ravel order.  The outer-product
// blocks are out[i, a] = sum_s w_s L[s, i] R[s, a] with (L, R) = (x, d1) and
// (h1, g2); a bias block takes L = 1.  Each workgroup sums a chunk of
// samples for a 16 x 64 tile of (i, a) into its own partial buffer, and the
// chunks are added in order afterwards.
// ---------------------------------------------------------------------------
constexpr int CT_I = 16, CT_A = 64, CT_S = 256;

template <class T>
__global__ void __launch_bounds__(256) ffnn_outer_partial(const int8_t* __restrict__ x, const T* __restrict__ L,
                                                          int nL, const T* __restrict__ R, int nR, int64_t ns,
                                                          const double* __restrict__ w, double* __restrict__ part,
                                                          int64_t plane) {
  __shared__ double sl[CT_S][CT_I + 1];
  __shared__ double sw[CT_S];
  const int ti = blockIdx.x, ta = blockIdx.y, chunk = blockIdx.z;
  const int i0 = ti * CT_I, a0 = ta * CT_A;
  const int64_t s0 = (int64_t)chunk * CT_S;
  const int64_t s1 = min<int64_t>(ns, s0 + CT_S);
  const int nsl = (int)(s1 - s0);
  for (int idx = threadIdx.x; idx < nsl * CT_I; idx += 256) {
    const int s = idx / CT_I, i = idx % CT_I;
    double val = 0.0;
    if (i0 + i < nL) {
      if (L == nullptr && x == nullptr) val = 1.0;
      else if (x) val = x[(s0 + s) * nL + i0 + i] > 0 ? 1.0 : -1.0;
      else val = (double)L[(s0 + s) * nL + i0 + i];
    }
    sl[s][i] = val;
  }
  for (int s = threadIdx.x; s < nsl; s += 256) sw[s] = w[s0 + s];
  __syncthreads();
  const int a = a0 + (threadIdx.x & 63);
  const int ig = threadIdx.x >> 6;  // 4 groups of 4 i each
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (a < nR) {
    for (int s = 0; s < nsl; ++s) {
      const double rv = (double)R[(s0 + s) * nR + a] * sw[s];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] += sl[s][ig * 4 + q] * rv;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + ig * 4 + q;
      if (i < nL) part[(int64_t)chunk * plane + (int64_t)i * nR + a] = acc[q];
    }
  }
}

// out[off + p] = sum_chunks part[chunk][p] - mean[off + p] * wsum, p < count
__global__ void ffnn_outer_reduce(const double* __restrict__ part, int nchunks, int64_t plane, int64_t count,
                                  const double* __restrict__ mean, int64_t off, const double* __restrict__ wsum,
                                  double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += part[(int64_t)c * plane + p];
  if (mean) s -= mean[off + p] * wsum[0];
  out[off + p] = s;
}

// wsum[0] = sum_s w[s] (one workgroup, fixed order per thread then tree)
__global__ void ffnn_wsum_kernel(const double* __restrict__ w, int64_t ns, double* __restrict__ wsum) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < ns; i += 256) s += w[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) wsum[0] = red[0];
}

// ---------------------------------------------------------------------------
// Structured centred NTK of the network:
//   (O O^T)_st = (hb2 + h1_s.h1_t)(g2_s.g2_t) + (hb1 + x_s.x_t)(d1_s.d1_t)
//   T_st = scale * [(O O^T)_st - v_s - v_t + c] + shift delta_st + offset
// v = O mean, c = |mean|^2.  Four FP64 Gram products with inner sizes H1,
// H2, N, H1 per 64x64 tile (4 waves of 32x32, v_mfma_f64_16x16x4).
// ---------------------------------------------------------------------------
constexpr int NF_T = 64, NF_K = 16, NF_LD = 17;

// One K slab (64 rows x NF_K) of both operands of an NTK tile into
// registers: thread tid holds elements tid + 256 it (row idx / 16, k idx % 16).
// Spins (pass over x) become +-1; entries past the edges are zero.
__device__ __forceinline__ void ntk_fetch(const double* __restrict__ src, const int8_t* __restrict__ x, int N,
                                          int K, int64_t ns, int64_t i0, int64_t j0, int k0, bool spins,
                                          double (&pa)[4], double (&pb)[4]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    const int idx = tid + it * 256;
    const int r = idx >> 4, kc = idx & 15;
    const int kk = k0 + kc;
    double va = 0.0, vb = 0.0;
    if (kk < K) {
      if (spins) {
        if (i0 + r < ns) va = x[(i0 + r) * N + kk] > 0 ? 1.0 : -1.0;
        if (j0 + r < ns) vb = x[(j0 + r) * N + kk] > 0 ? 1.0 : -1.0;
      } else {
        if (i0 + r < ns) va = src[(i0 + r) * K + kk];
        if (j0 + r < ns) vb = src[(j0 + r) * K + kk];
      }
    }
    pa[it] = va;
    pb[it] = vb;
  }
}

__global__ void __launch_bounds__(256) ffnn_ntk_kernel(const double* __restrict__ h1, const double* __restrict__ g2,
                                                       const int8_t* __restrict__ x, const double* __restrict__ d1,
                                                       int N, int H1, int H2, int64_t ns,
                                                       const double* __restrict__ vv, const double* __restrict__ cc,
                                                       double hb1, double hb2, double scale, double shift,
                                                       double offset, double* __restrict__ out, int64_t ldo,
                                                       int64_t n_tiles) {
  __shared__ double sA[NF_T * NF_LD];
  __shared__ double sB[NF_T * NF_LD];
  const int64_t G = gridDim.x;
  const int64_t bid = (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  if (bid >= n_tiles) return;
  int ti, tj;
  tri_supertile(bid, (ns + NF_T - 1) / NF_T, ti, tj);
  const int64_t i0 = (int64_t)ti * NF_T, j0 = (int64_t)tj * NF_T;
  const int tid = threadIdx.x, wid = tid / 64, lane = lane_id();
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  doublex4 acc[4][2][2];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[p][a][b] = (doublex4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int pass = 0; pass < 4; ++pass) {
    const int K = pass == 0 ? H1 : pass == 1 ? H2 : pass == 2 ? N : H1;
    const double* src = pass == 0 ? h1 : pass == 1 ? g2 : pass == 2 ? nullptr : d1;
    // the next K slab is loaded into registers while the current one is used
    double pa[4], pb[4];
    ntk_fetch(src, x, N, K, ns, i0, j0, 0, pass == 2, pa, pb);
    for (int k0 = 0; k0 < K; k0 += NF_K) {
      __syncthreads();
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + it * 256;
        const int r = idx >> 4, kc = idx & 15;
        sA[r * NF_LD + kc] = pa[it];
        sB[r * NF_LD + kc] = pb[it];
      }
      __syncthreads();
      if (k0 + NF_K < K) ntk_fetch(src, x, N, K, ns, i0, j0, k0 + NF_K, pass == 2, pa, pb);
#pragma unroll
      for (int ks = 0; ks < NF_K; ks += 4) {
        const int kq = ks + (lane >> 4);
        double av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = sA[(wm + a * 16 + (lane & 15)) * NF_LD + kq];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = sB[(wn + b * 16 + (lane & 15)) * NF_LD + kq];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[pass][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[pass][a][b], 0, 0, 0);
      }
    }
  }
  const double c = *cc;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t gi = i0 + wm + a * 16 + (lane >> 4) + 4 * e;
        const int64_t gj = j0 + wn + b * 16 + (lane & 15);
        if (gi < ns && gj <= gi) {
          const double oo = (hb2 + acc[0][a][b][e]) * acc[1][a][b][e] + (hb1 + acc[2][a][b][e]) * acc[3][a][b][e];
          double val = scale * (oo - vv[gi] - vv[gj] + c) + offset;
          if (gi == gj) val += shift;
          out[gi * ldo + gj] = val;
        }
      }
}

template <class T>
__global__ void ffnn_to_f64_kernel(const T* __restrict__ a, int64_t n, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (double)a[i];
}

// v_s = d1_s.mu_b1 + x_s.U1_s + g2_s.mu_b2 + h1_s.U2_s (one wave per sample); block 0 also writes c = |mu|^2
__global__ void __launch_bounds__(256) ffnn_ntk_vec_kernel(const double* __restrict__ h1, const double* __restrict__ g2,
                                                           const int8_t* __restrict__ x, const double* __restrict__ d1,
                                                           const double* __restrict__ U1, const double* __restrict__ U2,
                                                           int N, int H1, int H2, int64_t ns,
                                                           const double* __restrict__ mean, int64_t b1_off,
                                                           int64_t b2_off, int64_t np, double* __restrict__ vv,
                                                           double* __restrict__ cc) {
  const int wid = threadIdx.x >> 6, lane = lane_id();
  for (int64_t s = (int64_t)blockIdx.x * 4 + wid; s < ns; s += (int64_t)gridDim.x * 4) {
    double acc = 0.0;
    for (int a = lane; a < H1; a += 64) {
      acc += h1[s * H1 + a] * U2[s * H1 + a];
      if (b1_off >= 0) acc += d1[s * H1 + a] * mean[b1_off + a];
    }
    for (int i = lane; i < N; i += 64) acc += (x[s * N + i] > 0 ? 1.0 : -1.0) * U1[s * N + i];
    if (b2_off >= 0)
      for (int k = lane; k < H2; k += 64) acc += g2[s * H2 + k] * mean[b2_off + k];
    acc = wave_sum(acc);
    if (lane == 0) vv[s] = acc;
  }
  if (blockIdx.x == 0) {
    __shared__ double red[4];
    double q = 0.0;
    for (int64_t p = threadIdx.x; p < np; p += 256) q += mean[p] * mean[p];
    q = wave_sum(q);
    if (lane == 0) red[wid] = q;
    __syncthreads();
    if (threadIdx.x == 0) cc[0] = red[0] + red[1] + red[2] + red[3];
  }
}

// ---------------------------------------------------------------------------
// Centred, rescaled Jacobian (same layouts and output kinds as the RBM's).
// ---------------------------------------------------------------------------
template <class T>
__device__ inline double ffnn_o(const nka_ffnn& m, const int8_t* x, const T* h1, const T* d1, const T* g2,
                                int64_t s, int64_t p) {
  const int N = m.n_visible, H1 = m.n_hidden1, H2 = m.n_hidden2;
  if (m.b1_off >= 0 && p >= m.b1_off && p < m.b1_off + H1) return (double)d1[s * H1 + (p - m.b1_off)];
  if (p >= m.w1_off && p < m.w1_off + (int64_t)N * H1) {
    const int64_t q = p - m.w1_off;
    const int i = (int)(q / H1), a = (int)(q % H1);
    return (x[s * N + i] > 0 ? 1.0 : -1.0) * (double)d1[s * H1 + a];
  }
  if (m.b2_off >= 0 && p >= m.b2_off && p < m.b2_off + H2) return (double)g2[s * H2 + (p - m.b2_off)];
  const int64_t q = p - m.w2_off;
  const int a = (int)(q / H2), k = (int)(q % H2);
  return (double)h1[s * H1 + a] * (double)g2[s * H2 + k];
}

template <class T>
__global__ void ffnn_jacobian_kernel(nka_ffnn m, const int8_t* __restrict__ x, const T* __restrict__ h1,
                                     const T* __restrict__ d1, const T* __restrict__ g2, int64_t ns, int64_t np,
                                     const double* __restrict__ mean, double scale, int layout, int kind,
                                     void* __restrict__ oa, void* __restrict__ ob, int64_t ldo) {
  const int64_t total = ns * np;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t s, p;
    if (layout == NKA_LAYOUT_PS) {
      p = idx / ns;
      s = idx % ns;
    } else {
      s = idx / np;
      p = idx % np;
    }
    double o = ffnn_o<T>(m, x, h1, d1, g2, s, p);
    if (mean) o -= mean[p];
    o *= scale;
    const int64_t off = layout == NKA_LAYOUT_PS ? p * ldo + s : s * ldo + p;
    if (kind == NKA_OUT_F64) {
      ((double*)oa)[off] = o;
    } else if (kind == NKA_OUT_F32) {
      ((float*)oa)[off] = (float)o;
    } else {
      const float f = (float)o;
      const uint16_t hi = bf16_rne(f);
      const float rem = f - __uint_as_float((uint32_t)hi << 16);
      ((uint16_t*)oa)[off] = hi;
      ((uint16_t*)ob)[off] = bf16_rne(rem);
    }
  }
}

// ---------------------------------------------------------------------------
// Metropolis sampler: 32 chains per workgroup in lock step.  Thread t holds
// theta1 of chain t / 8 for the units a = t % 8 + 8 j in registers; the chain
// spins live in LDS.  Each proposal: wave w draws the proposals of chains
// 8w .. 8w+7 (same rules and Philox stream layout as the RBM sampler:
// counter q, chain id, TAG_METROPOLIS; x word picks the site / cluster, the
// uniform of the acceptance comes from the same draw); theta1' and
// h1' = log_cosh(theta1') are formed, the second layer gives log psi'; a
// chain accepts when u < exp(machine_pow (log psi' - log psi) + corr).
// ---------------------------------------------------------------------------
struct FfSweepArgs {
  int rule;
  const int32_t* clusters;
  int n_clusters;
  const uint32_t* weights;      // exchange: integer weight per cluster (NULL: uniform)
  int C, sweep_size, n_discard, chain_length;
  double machine_pow;
  uint32_t k0, k1;
  uint64_t counter;
  const uint64_t* counter_dev;
};

constexpr int FF_MAXCL = 16;  // cluster slots per lane: n_clusters <= 1024

template <class T, int HPT>
__global__ void __launch_bounds__(FF_THREADS) ffnn_metropolis_kernel(nka_ffnn m, FfSweepArgs sa,
                                                                    int8_t* __restrict__ sigma, T* __restrict__ log_prob,
                                                                    int8_t* __restrict__ samples,
                                                                    T* __restrict__ lp_samples,
                                                                    int64_t* __restrict__ n_accepted) {
  extern __shared__ unsigned char smem[];
  const FfnnView<T> v = ffnn_view<T>(m);
  const int N = v.N, H1 = v.H1, H2 = v.H2;
  const int lda = lda_of(H1);
  T* hA = (T*)smem;
  T* part = hA + FF_ROWS * lda;
  T* rowsum = part + FF_ROWS * FF_MAXBLK;
  T* lp = rowsum + FF_ROWS;                       // current log prob per chain
  T* pcorr = lp + FF_ROWS;                        // proposal correction
  T* pu = pcorr + FF_ROWS;                        // acceptance uniform
  int* psite = (int*)(pu + FF_ROWS);              // 2 sites per chain
  int* pnf = psite + 2 * FF_ROWS;                 // flipped sites (0: no move)
  int* pacc = pnf + FF_ROWS;                      // accepted flag
  int* nacc = pacc + FF_ROWS;
  int8_t* xs = (int8_t*)(nacc + FF_ROWS);         // 32 x N spins
  // exchange rule with equal weights: each chain's differing clusters as a
  // bit set (word w, bit b: cluster 64 w + b) and their count, the clusters,
  // and for every site the clusters that contain it (other site, cluster)
  const bool exch = sa.rule == NKA_RULE_EXCHANGE;
  const bool fast_exch = exch && !sa.weights;
  const int ncl = fast_exch ? sa.n_clusters : 0;
  const int NWc = (ncl + 63) / 64;
  unsigned char* xb = smem + ffpad((int64_t)((unsigned char*)xs - smem) + (int64_t)FF_ROWS * N);
  uint64_t* dmask = (uint64_t*)xb;                                   // FF_ROWS x NWc
  int* cls = (int*)(xb + ffpad((int64_t)FF_ROWS * NWc * 8));         // 2 ncl
  int* adj_off = cls + 2 * ncl;                                      // N + 1
  int* adj_pos = adj_off + (N + 1);                                  // N
  int* adj_t = adj_pos + N;                                          // 2 ncl
  int* adj_c = adj_t + 2 * ncl;                                      // 2 ncl
  int* ndiff = adj_c + 2 * ncl;                                      // FF_ROWS
  int* pnew = ndiff + FF_ROWS;                                       // FF_ROWS
  T* rbuf = sizeof(T) == 4 ? (T*)(smem + ffpad((int64_t)((unsigned char*)(pnew + FF_ROWS) - smem))) : nullptr;
  const int tid = threadIdx.x, wid = tid >> 6, lane = lane_id();
  const int c0 = blockIdx.x * FF_ROWS;
  const int cr = tid / FF_GROUPS, g = tid % FF_GROUPS;   // chain row and unit group of this thread
  const int chain = c0 + cr;

  for (int idx = tid; idx < FF_ROWS * N; idx += FF_THREADS) {
    const int r = idx / N, i = idx % N;
    xs[idx] = (c0 + r < sa.C) ? sigma[(int64_t)(c0 + r) * N + i] : (int8_t)1;
  }
  if (tid < FF_ROWS) nacc[tid] = 0;
  if (fast_exch) {
    for (int i = tid; i <= N; i += FF_THREADS) adj_off[i] = 0;
    for (int i = tid; i < 2 * ncl; i += FF_THREADS) cls[i] = sa.clusters[i];
    __syncthreads();
    for (int c = tid; c < ncl; c += FF_THREADS) {
      atomicAdd(&adj_off[cls[2 * c] + 1], 1);
      atomicAdd(&adj_off[cls[2 * c + 1] + 1], 1);
    }
    __syncthreads();
    if (tid == 0)
      for (int i = 0; i < N; ++i) adj_off[i + 1] += adj_off[i];
    __syncthreads();
    for (int i = tid; i < N; i += FF_THREADS) adj_pos[i] = adj_off[i];
    for (int idx = tid; idx < FF_ROWS * NWc; idx += FF_THREADS) {
      const int r = idx / NWc, w = idx % NWc;
      uint64_t bits = 0;
      for (int b = 0; b < 64; ++b) {
        const int c = 64 * w + b;
        if (c < ncl && xs[r * N + cls[2 * c]] != xs[r * N + cls[2 * c + 1]]) bits |= 1ull << b;
      }
      dmask[idx] = bits;
    }
    __syncthreads();
    for (int c = tid; c < ncl; c += FF_THREADS) {
      const int u = cls[2 * c], w = cls[2 * c + 1];
      int at = atomicAdd(&adj_pos[u], 1);
      adj_t[at] = w;
      adj_c[at] = c;
      at = atomicAdd(&adj_pos[w], 1);
      adj_t[at] = u;
      adj_c[at] = c;
    }
    if (tid < FF_ROWS) {
      int cnt = 0;
      for (int w = 0; w < NWc; ++w) cnt += __popcll(dmask[tid * NWc + w]);
      ndiff[tid] = cnt;
    }
  }
  __syncthreads();
  T th[HPT];
#pragma unroll
  for (int j = 0; j < HPT; ++j) {
    const int a = g + FF_GROUPS * j;
    th[j] = a < H1 ? theta1_full<T>(v, xs + cr * N, a) : (T)0;
    if (a < H1) hA[cr * lda + a] = log_cosh<T>(th[j]);
  }
  __syncthreads();
  const T* W2 = v.W2;
  const T* b2 = v.b2;
  auto epi = [=](int, int c, T z) { return log_cosh<T>(z + (b2 ? b2[c] : (T)0)); };
  rows_gemm(hA, lda, H1, H2, W2, H2, 1, epi, part, rowsum, rbuf);
  const T pw = (T)sa.machine_pow;
  if (tid < FF_ROWS) lp[tid] = pw * rowsum[tid];

  // cluster slots of this lane (weighted exchange rule): cl = lane + 64 r
  const int RS = exch ? (sa.n_clusters + 63) / 64 : 0;
  int cu[FF_MAXCL], cv[FF_MAXCL];
#pragma unroll
  for (int r = 0; r < FF_MAXCL; ++r) {
    const int cl = lane + 64 * r;
    const bool ok = r < RS && cl < sa.n_clusters;
    cu[r] = ok ? sa.clusters[2 * cl] : -1;
    cv[r] = ok ? sa.clusters[2 * cl + 1] : -1;
  }
  uint64_t q = sa.counter;
  if (sa.counter_dev) {
    const uint64_t cd = *sa.counter_dev;
    q = ((uint64_t)uniform_value<uint32_t>((uint32_t)(cd >> 32)) << 32) |
        (uint64_t)uniform_value<uint32_t>((uint32_t)cd);
  }
  __syncthreads();
  const int n_sweeps = sa.n_discard + sa.chain_length;
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int p = 0; p < sa.sweep_size; ++p, ++q) {
      if (!exch || fast_exch) {
        // proposals of chains 8w .. 8w+7 by lanes 0..7 of wavefront w.  An
        // exchange picks the k-th set bit of the chain's differing-cluster
        // set; swapping the two (different) spins of that cluster flips the
        // state of every other cluster at either site, so the count after the
        // move follows from the sites' clusters alone
        if (lane < FF_ROWS / 4) {
          const int r = wid * (FF_ROWS / 4) + lane;
          const int ch = c0 + r;
          const u32x4 rnd = philox4x32_10({(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)ch, TAG_METROPOLIS},
                                          sa.k0, sa.k1);
          const int8_t* x = xs + r * N;
          int s0 = 0, s1 = 0, nf = 1, nn = 0;
          T corr = (T)0;
          if (!exch) {
            s0 = (int)bounded(rnd.x, (uint32_t)N);
          } else {
            const int nd = ndiff[r];
            nf = 0;
            if (nd > 0) {
              int k = (int)bounded(rnd.x, (uint32_t)nd);
              const uint64_t* dm = dmask + r * NWc;
              int w = 0;
              uint64_t word = dm[0];
              int cnt = __popcll(word);
              while (k >= cnt && w + 1 < NWc) {
                k -= cnt;
                ++w;
                word = dm[w];
                cnt = __popcll(word);
              }
              if (k < cnt) {
                const int c = 64 * w + kth_set_bit(word, k);
                s0 = cls[2 * c];
                s1 = cls[2 * c + 1];
                nf = 2;
                const int8_t x0 = x[s0], x1 = x[s1];
                int delta = 0;
                for (int e = adj_off[s0]; e < adj_off[s0 + 1]; ++e) {
                  const int t = adj_t[e];
                  if (t != s0 && t != s1) delta += x[t] != x0 ? -1 : 1;
                }
                for (int e = adj_off[s1]; e < adj_off[s1 + 1]; ++e) {
                  const int t = adj_t[e];
                  if (t != s0 && t != s1) delta += x[t] != x1 ? -1 : 1;
                }
                nn = nd + delta;
                corr = log_t<T>((T)nd) - log_t<T>((T)nn);
              }
            }
          }
          psite[2 * r] = s0;
          psite[2 * r + 1] = s1;
          pnf[r] = nf;
          pcorr[r] = corr;
          pu[r] = uniform01<T>(rnd);
          pnew[r] = nn;
        }
      } else
      // weighted exchange: proposals of chains 8w .. 8w+7, one at a time per wavefront
      for (int cc = 0; cc < FF_ROWS / 4; ++cc) {
        const int r = wid * (FF_ROWS / 4) + cc;
        const int ch = c0 + r;
        const u32x4 rnd = philox4x32_10({(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)ch, TAG_METROPOLIS},
                                        sa.k0, sa.k1);
        const int8_t* x = xs + r * N;
        int s0 = 0, s1 = 0, nf = 1;
        T corr = (T)0;
        if (!exch) {
          s0 = (int)bounded(uniform_value<uint32_t>(rnd.x), (uint32_t)N);
        } else if (sa.weights) {
          // weighted choice among the differing clusters (as the RBM sampler)
          uint32_t wd[FF_MAXCL];
          uint32_t tot = 0;
#pragma unroll
          for (int rr = 0; rr < FF_MAXCL; ++rr) {
            const bool d = rr < RS && cu[rr] >= 0 && x[cu[rr]] != x[cv[rr]];
            wd[rr] = d ? sa.weights[lane + 64 * rr] : 0u;
            if (rr < RS) tot += wave_sum_u32(wd[rr]);
          }
          tot = uniform_value<uint32_t>(tot);
          if (tot == 0) {
            nf = 0;
          } else {
            const uint32_t k = bounded(uniform_value<uint32_t>(rnd.x), tot);
            int sel = uniform_value<int>(weighted_pick<FF_MAXCL>(wd, RS, k));
            if (sel < 0 || sel >= sa.n_clusters) sel = 0;  // unreachable: k < tot
            cluster_sites<FF_MAXCL>(cu, cv, sel, s0, s1);
            nf = 2;
            uint32_t tot_new = 0;
#pragma unroll
            for (int rr = 0; rr < FF_MAXCL; ++rr) {
              if (rr < RS) {
                bool d = false;
                if (cu[rr] >= 0) {
                  const int ua = cu[rr], ub = cv[rr];
                  const int8_t xa = ua == s0 ? x[s1] : (ua == s1 ? x[s0] : x[ua]);
                  const int8_t xb = ub == s0 ? x[s1] : (ub == s1 ? x[s0] : x[ub]);
                  d = xa != xb;
                }
                tot_new += wave_sum_u32(d ? sa.weights[lane + 64 * rr] : 0u);
              }
            }
            tot_new = uniform_value<uint32_t>(tot_new);
            corr = uniform_value<T>(log_t<T>((T)tot) - log_t<T>((T)tot_new));
          }
        } else {
          uint64_t bal[FF_MAXCL];
          int n_diff = 0;
#pragma unroll
          for (int rr = 0; rr < FF_MAXCL; ++rr) {
            const bool d = rr < RS && cu[rr] >= 0 && x[cu[rr]] != x[cv[rr]];
            bal[rr] = __ballot(d);
            n_diff += __popcll(bal[rr]);
          }
          n_diff = uniform_value<int>(n_diff);
          if (n_diff == 0) {
            nf = 0;
          } else {
            int k = (int)bounded(uniform_value<uint32_t>(rnd.x), (uint32_t)n_diff);
            int sel = -1;
#pragma unroll
            for (int rr = 0; rr < FF_MAXCL; ++rr) {
              if (sel < 0 && rr < RS) {
                const int cnt = __popcll(bal[rr]);
                if (k < cnt) sel = rr * 64 + kth_set_bit(bal[rr], k);
                else k -= cnt;
              }
            }
            sel = uniform_value<int>(sel);
            if (sel < 0 || sel >= sa.n_clusters) sel = 0;  // unreachable: k < n_diff
            cluster_sites<FF_MAXCL>(cu, cv, sel, s0, s1);
            nf = 2;
            int n_new = 0;
#pragma unroll
            for (int rr = 0; rr < FF_MAXCL; ++rr) {
              bool d = false;
              if (rr < RS && cu[rr] >= 0) {
                const int ua = cu[rr], ub = cv[rr];
                const int8_t xa = ua == s0 ? x[s1] : (ua == s1 ? x[s0] : x[ua]);
                const int8_t xb = ub == s0 ? x[s1] : (ub == s1 ? x[s0] : x[ub]);
                d = xa != xb;
              }
              n_new += __popcll(__ballot(d));
            }
            n_new = uniform_value<int>(n_new);
            corr = uniform_value<T>(log_t<T>((T)n_diff) - log_t<T>((T)n_new));
          }
        }
        if (lane == 0) {
          psite[2 * r] = s0;
          psite[2 * r + 1] = s1;
          pnf[r] = nf;
          pcorr[r] = corr;
          pu[r] = uniform01<T>(rnd);
        }
      }
      __syncthreads();
      // theta1' and h1' of every chain
      T thn[HPT];
      {
        // at most two flipped sites; their W1 rows are read with every load
        // of the thread's units in flight (no loop over the flips per unit)
        const int nf = pnf[cr];
        const int8_t* x = xs + cr * N;
        const int s0 = nf > 0 ? psite[2 * cr] : 0;
        const int s1 = nf > 1 ? psite[2 * cr + 1] : 0;
        // x'_site = -x_site: theta1 gains (x' - x) W1 = -2 x W1
        const T c0 = x[s0] > 0 ? (T)-2 : (T)2;
        const T c1 = x[s1] > 0 ? (T)-2 : (T)2;
        const T* w0 = v.W1 + (int64_t)s0 * H1;
        const T* w1 = v.W1 + (int64_t)s1 * H1;
#pragma unroll
        for (int j = 0; j < HPT; ++j) {
          const int a = g + FF_GROUPS * j;
          T t = th[j];
          if (a < H1) {
            const T u0 = w0[a], u1 = w1[a];
            if (nf > 0) t += c0 * u0;
            if (nf > 1) t += c1 * u1;
            hA[cr * lda + a] = log_cosh<T>(t);
          }
          thn[j] = t;
        }
      }
      __syncthreads();
      rows_gemm(hA, lda, H1, H2, W2, H2, 1, epi, part, rowsum, rbuf);
      if (tid < FF_ROWS) {
        const int r = tid;
        int acc = 0;
        if (pnf[r] > 0 && c0 + r < sa.C) {
          const T lpn = pw * rowsum[r];
          acc = pu[r] < exp_t<T>(lpn - lp[r] + pcorr[r]);
          if (acc) {
            lp[r] = lpn;
            for (int f = 0; f < pnf[r]; ++f) {
              const int site = psite[2 * r + f];
              xs[r * N + site] = (int8_t)(-xs[r * N + site]);
            }
            nacc[r] += 1;
            if (fast_exch) ndiff[r] = pnew[r];
          }
        }
        pacc[r] = acc;
      }
      __syncthreads();
      if (pacc[cr]) {
#pragma unroll
        for (int j = 0; j < HPT; ++j) th[j] = thn[j];
        if (fast_exch) {
          // every other cluster at the two exchanged sites changes state; the
          // chain's FF_GROUPS threads share its clusters
          const int s0 = psite[2 * cr], s1 = psite[2 * cr + 1];
          unsigned long long* dm = (unsigned long long*)(dmask + cr * NWc);
          for (int e = adj_off[s0] + g; e < adj_off[s0 + 1]; e += FF_GROUPS) {
            const int t = adj_t[e], c = adj_c[e];
            if (t != s0 && t != s1) atomicXor(&dm[c >> 6], 1ull << (c & 63));
          }
          for (int e = adj_off[s1] + g; e < adj_off[s1 + 1]; e += FF_GROUPS) {
            const int t = adj_t[e], c = adj_c[e];
            if (t != s0 && t != s1) atomicXor(&dm[c >> 6], 1ull << (c & 63));
          }
        }
      }
      if (fast_exch) __syncthreads();  // the next proposals read the updated sets
    }
    if (sw >= sa.n_discard) {
      const int t = sw - sa.n_discard;
      for (int idx = tid; idx < FF_ROWS * N; idx += FF_THREADS) {
        const int r = idx / N, i = idx % N;
        if (c0 + r < sa.C) samples[((int64_t)(c0 + r) * sa.chain_length + t) * N + i] = xs[idx];
      }
      if (lp_samples && tid < FF_ROWS && c0 + tid < sa.C)
        lp_samples[(int64_t)(c0 + tid) * sa.chain_length + t] = lp[tid];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < FF_ROWS * N; idx += FF_THREADS) {
    const int r = idx / N, i = idx % N;
    if (c0 + r < sa.C) sigma[(int64_t)(c0 + r) * N + i] = xs[idx];
  }
  if (tid < FF_ROWS && c0 + tid < sa.C) {
    log_prob[c0 + tid] = lp[tid];
    n_accepted[c0 + tid] += nacc[tid];
  }
  (void)chain;
}

// ncl: clusters of an exchange rule with equal weights (0 otherwise)
template <class T>
inline size_t metropolis_lds(const nka_ffnn& m, int ncl) {
  const int64_t nw = (ncl + 63) / 64, N = m.n_visible;
  return ffpad((size_t)FF_ROWS * lda_of(m.n_hidden1) * sizeof(T)) +
         ffpad((size_t)FF_ROWS * FF_MAXBLK * sizeof(T)) + ffpad((size_t)4 * FF_ROWS * sizeof(T)) +
         ffpad((size_t)5 * FF_ROWS * sizeof(int)) + ffpad((size_t)FF_ROWS * m.n_visible) +
         ffpad(FF_ROWS * nw * 8) + ffpad(((size_t)6 * ncl + 2 * N + 1 + 2 * FF_ROWS) * sizeof(int)) +
         (sizeof(T) == 4 ? split_scratch_bytes(m.n_hidden2) : 0);
}

// Dynamic LDS above the default 64 KB needs the function attribute.
template <class K>
static void allow_lds(K kernel, int64_t bytes) {
  if (bytes > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <class T, int HPT>
static void launch_ffnn_sweep(const nka_ffnn& m, const FfSweepArgs& sa, int8_t* sigma, void* lp, int8_t* smp,
                              void* lps, int64_t* nacc, hipStream_t s) {
  const unsigned g = (unsigned)((sa.C + FF_ROWS - 1) / FF_ROWS);
  const int ncl = (sa.rule == NKA_RULE_EXCHANGE && !sa.weights) ? sa.n_clusters : 0;
  const size_t lds = metropolis_lds<T>(m, ncl);
  allow_lds(ffnn_metropolis_kernel<T, HPT>, (int64_t)lds);
  hipLaunchKernelGGL((ffnn_metropolis_kernel<T, HPT>), dim3(g), dim3(FF_THREADS), lds, s, m, sa, sigma, (T*)lp,
                     smp, (T*)lps, nacc);
}

template <class T>
static void dispatch_ffnn_sweep(const nka_ffnn& m, const FfSweepArgs& sa, int8_t* sigma, void* lp, int8_t* smp,
                                void* lps, int64_t* nacc, hipStream_t s) {
  const int need = (m.n_hidden1 + FF_GROUPS - 1) / FF_GROUPS;
  if (need <= 8) launch_ffnn_sweep<T, 8>(m, sa, sigma, lp, smp, lps, nacc, s);
  else if (need <= 16) launch_ffnn_sweep<T, 16>(m, sa, sigma, lp, smp, lps, nacc, s);
  else if (need <= 32) launch_ffnn_sweep<T, 32>(m, sa, sigma, lp, smp, lps, nacc, s);
  else if (need <= 56) launch_ffnn_sweep<T, 56>(m, sa, sigma, lp, smp, lps, nacc, s);
  else if (need <= 64) launch_ffnn_sweep<T, 64>(m, sa, sigma, lp, smp, lps, nacc, s);
  else launch_ffnn_sweep<T, 128>(m, sa, sigma, lp, smp, lps, nacc, s);
}

static int check_model(const nka_ffnn* m, const char* who) {
  NKA_CHECK_ARG(m && m->params, "%s: null model", who);
  NKA_CHECK_ARG(m->n_visible > 0 && m->n_hidden1 > 0 && m->n_hidden2 > 0, "%s: bad sizes", who);
  NKA_CHECK_ARG(m->n_hidden1 <= FF_MAXH && m->n_hidden2 <= FF_MAXH, "%s: hidden layers larger than %d", who,
                FF_MAXH);
  NKA_CHECK_ARG(m->dtype == NKA_F32 || m->dtype == NKA_F64, "%s: bad dtype", who);
  const int64_t lds = m->dtype == NKA_F32
                          ? std::max<int64_t>(std::max<int64_t>(forward_lds<float>(*m), backprop_lds<float>(*m)),
                                              metropolis_lds<float>(*m, 0))
                          : std::max<int64_t>(std::max<int64_t>(forward_lds<double>(*m), backprop_lds<double>(*m)),
                                              metropolis_lds<double>(*m, 0));
  NKA_CHECK_ARG(lds <= FF_LDS_MAX, "%s: layer sizes need %lld bytes of LDS per workgroup (max %lld)", who,
                (long long)lds, (long long)FF_LDS_MAX);
  return NKA_OK;
}

inline int64_t ffnn_np(const nka_ffnn& m) {
  return (m.b1_off >= 0 ? m.n_hidden1 : 0) + (int64_t)m.n_visible * m.n_hidden1 +
         (m.b2_off >= 0 ? m.n_hidden2 : 0) + (int64_t)m.n_hidden1 * m.n_hidden2;
}

template <class T>
static void launch_forward0(const nka_ffnn& m, const int8_t* x, int64_t rows, T* lp, T* th, T* g2, hipStream_t s) {
  if (rows <= 0) return;
  const unsigned g = (unsigned)((rows + FF_ROWS - 1) / FF_ROWS);
  allow_lds(ffnn_forward_kernel<T, 0>, (int64_t)forward_lds<T>(m));
  hipLaunchKernelGGL((ffnn_forward_kernel<T, 0>), dim3(g), dim3(FF_THREADS), forward_lds<T>(m), s, m, x,
                     (const int8_t*)nullptr, 1, (const T*)nullptr, rows, lp, th, g2, (const int*)nullptr,
                     (const int*)nullptr);
}

static int op_rows(int op_kind, const void* op, int N, int* K, int* mel_dtype) {
  if (op_kind == NKA_OP_ISING) {
    const nka_ising* o = (const nka_ising*)op;
    *K = o->h == 0.0 ? 1 : N + 1;
    *mel_dtype = o->mel_dtype;
  } else if (op_kind == NKA_OP_LOCAL) {
    const nka_localop* o = (const nka_localop*)op;
    *K = o->max_conn_size;
    *mel_dtype = o->mel_dtype;
  } else {
    return NKA_ERR_ARG;
  }
  return NKA_OK;
}

// bytes per sample of the local-energy workspace
static int64_t eloc_bytes_per_sample(const nka_ffnn& m, int K) {
  const int64_t ts = m.dtype == NKA_F32 ? 4 : 8;
  return ffpad((int64_t)K * m.n_visible) + (int64_t)K * 8 + (int64_t)K * ts + (int64_t)m.n_hidden1 * ts + ts +
         (int64_t)K * 4 + 4;
}

template <class T>
static int ffnn_eloc_impl(const nka_ffnn& m, int op_kind, const void* op, const int8_t* sigma, int64_t n,
                          double* eloc, T* log_psi, unsigned char* ws, int64_t ws_bytes, hipStream_t s) {
  int K = 0, md = 0;
  if (op_rows(op_kind, op, m.n_visible, &K, &md) != NKA_OK) {
    set_error("nka_ffnn_local_energy: unknown operator kind %d", op_kind);
    return NKA_ERR_ARG;
  }
  const int N = m.n_visible, H1 = m.n_hidden1;
  const int64_t per = eloc_bytes_per_sample(m, K);
  const int64_t chunk = std::min<int64_t>(n, (ws_bytes - 4096) / per);
  NKA_CHECK_ARG(chunk >= 1, "nka_ffnn_local_energy: workspace too small");
  unsigned char* p = ws;
  auto take = [&](int64_t bytes) {
    unsigned char* q = p;
    p += ffpad(bytes);
    return q;
  };
  int8_t* xp = (int8_t*)take(chunk * K * N);
  void* mels = take(chunk * K * 8);
  T* lp = (T*)take(chunk * K * sizeof(T));
  T* th = (T*)take(chunk * H1 * sizeof(T));
  T* lp0own = (T*)take(chunk * sizeof(T));
  int* ridx = (int*)take(chunk * K * (int64_t)sizeof(int));
  int* rcount = (int*)take(16);
  NKA_CHECK_ARG(chunk * (int64_t)K < (1LL << 31), "nka_ffnn_local_energy: too many connected rows per chunk");
  for (int64_t s0 = 0; s0 < n; s0 += chunk) {
    const int64_t cs = std::min<int64_t>(chunk, n - s0);
    const int8_t* x = sigma + s0 * N;
    T* lp0 = log_psi ? log_psi + s0 : lp0own;
    int st;
    if (op_kind == NKA_OP_ISING) {
      const nka_ising* o = (const nka_ising*)op;
      st = nka_ising_conn(x, cs, N, o->edges, o->n_edges, o->h, o->J, o->mel_dtype, K, xp, mels, s);
    } else {
      st = nka_localop_conn(x, cs, N, (const nka_localop*)op, xp, mels, nullptr, s);
    }
    if (st) return st;
    launch_forward0<T>(m, x, cs, lp0, th, nullptr, s);
    const int64_t rows = cs * K;
    const unsigned g = (unsigned)((rows + FF_ROWS - 1) / FF_ROWS);
    hipLaunchKernelGGL(ffnn_zero_count_kernel, dim3(1), dim3(64), 0, s, rcount);
    const unsigned gc = (unsigned)std::min<int64_t>((rows + 255) / 256, 8192);
    if (md == NKA_F32)
      hipLaunchKernelGGL((ffnn_compact_rows_kernel<float>), dim3(gc), dim3(256), 0, s, (const float*)mels, rows,
                         ridx, rcount);
    else
      hipLaunchKernelGGL((ffnn_compact_rows_kernel<double>), dim3(gc), dim3(256), 0, s, (const double*)mels, rows,
                         ridx, rcount);
    allow_lds(ffnn_forward_kernel<T, 1>, (int64_t)forward_lds<T>(m));
    hipLaunchKernelGGL((ffnn_forward_kernel<T, 1>), dim3(g), dim3(FF_THREADS), forward_lds<T>(m), s, m, xp, x, K,
                       (const T*)th, rows, lp, (T*)nullptr, (T*)nullptr, (const int*)ridx, (const int*)rcount);
    const unsigned gr = (unsigned)((cs + 3) / 4);
    if (md == NKA_F32)
      hipLaunchKernelGGL((ffnn_eloc_reduce_kernel<T, float>), dim3(gr), dim3(256), 0, s, (const float*)mels,
                         (const T*)lp, (const T*)lp0, cs, K, eloc + s0);
    else
      hipLaunchKernelGGL((ffnn_eloc_reduce_kernel<T, double>), dim3(gr), dim3(256), 0, s, (const double*)mels,
                         (const T*)lp, (const T*)lp0, cs, K, eloc + s0);
  }
  return NKA_OK;
}

template <class T>
static void contract_one(const nka_ffnn& m, const int8_t* sigma, const T* h1, const T* d1, const T* g2, int64_t ns,
                         const double* w, const double* mean, double* out, double* part, double* wsum,
                         hipStream_t s) {
  const int N = m.n_visible, H1 = m.n_hidden1, H2 = m.n_hidden2;
  const int nch = (int)((ns + CT_S - 1) / CT_S);
  hipLaunchKernelGGL(ffnn_wsum_kernel, dim3(1), dim3(256), 0, s, w, ns, wsum);
  struct Blk {
    int64_t off;
    const int8_t* x;
    const T* L;
    int nL;
    const T* R;
    int nR;
  };
  Blk blks[4];
  int nb = 0;
  if (m.b1_off >= 0) blks[nb++] = {m.b1_off, nullptr, nullptr, 1, d1, H1};
  blks[nb++] = {m.w1_off, sigma, nullptr, N, d1, H1};
  if (m.b2_off >= 0) blks[nb++] = {m.b2_off, nullptr, nullptr, 1, g2, H2};
  blks[nb++] = {m.w2_off, nullptr, h1, H1, g2, H2};
  for (int b = 0; b < nb; ++b) {
    const Blk& k = blks[b];
    const int64_t plane = (int64_t)k.nL * k.nR;
    dim3 grid((unsigned)((k.nL + CT_I - 1) / CT_I), (unsigned)((k.nR + CT_A - 1) / CT_A), (unsigned)nch);
    hipLaunchKernelGGL((ffnn_outer_partial<T>), grid, dim3(256), 0, s, k.x, k.L, k.nL, k.R, k.nR, ns, w, part,
                       plane);
    hipLaunchKernelGGL(ffnn_outer_reduce, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s,
                       (const double*)part, nch, plane, plane, mean, k.off, (const double*)wsum, out);
  }
}

// y = alpha (raw - mean.v) + beta y (nka_ntk.hip)
void launch_ntk_split(const float* t, int64_t ns, int M, int ldt, uint16_t* hi, uint16_t* lo, hipStream_t s);
void launch_ntk_pack_spins(const int8_t* sigma, int64_t ns, int N, int nw, uint32_t* bits, hipStream_t s);
bool launch_gram_ffnn(const uint16_t* h1h, const uint16_t* h1l, int64_t K0, int64_t ld0, const uint16_t* g2h,
                      const uint16_t* g2l, int64_t K1, int64_t ld1, const uint16_t* d1h, const uint16_t* d1l,
                      int64_t K2, int64_t ld2, int64_t n, const uint32_t* bits, int nw, int N, const double* v,
                      const double* c, double hb1, double hb2, double scale, double shift, double offset,
                      double* out, int64_t ldo, hipStream_t s);
void launch_jvp_finish(const double* raw, int64_t ns, const double* mean, const double* v, int64_t np,
                       double* dot_slot, double alpha, double beta, double* y, hipStream_t s);

}  // namespace nka

using namespace nka;

extern "C" {

int nka_ffnn_log_psi(const nka_ffnn* m, const int8_t* sigma, int64_t n, void* log_psi, void* stream) {
  if (int st = check_model(m, "nka_ffnn_log_psi")) return st;
  NKA_CHECK_ARG(n >= 0 && (n == 0 || (sigma && log_psi)), "nka_ffnn_log_psi: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  if (m->dtype == NKA_F32)
    launch_forward0<float>(*m, sigma, n, (float*)log_psi, nullptr, nullptr, s);
  else
    launch_forward0<double>(*m, sigma, n, (double*)log_psi, nullptr, nullptr, s);
  NKA_CHECK_LAUNCH("nka_ffnn_log_psi");
  return NKA_OK;
}

int64_t nka_ffnn_local_energy_workspace(const nka_ffnn* m, int op_kind, const void* op, int64_t n) {
  if (!m || !op || n <= 0) return 0;
  int K = 0, md = 0;
  if (op_rows(op_kind, op, m->n_visible, &K, &md) != NKA_OK || K <= 0) return 0;
  const int64_t per = eloc_bytes_per_sample(*m, K);
  const int64_t cap = (int64_t)512 << 20;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, cap / per));
  return chunk * per + 4096;
}

int nka_ffnn_local_energy(const nka_ffnn* m, int op_kind, const void* op, const int8_t* sigma, int64_t n,
                          double* eloc, void* log_psi, void* workspace, int64_t workspace_bytes, void* stream) {
  if (int st = check_model(m, "nka_ffnn_local_energy")) return st;
  NKA_CHECK_ARG(op && n >= 0 && (n == 0 || (sigma && eloc && workspace)), "nka_ffnn_local_energy: bad arguments");
  if (n == 0) return NKA_OK;
  hipStream_t s = (hipStream_t)stream;
  int st;
  if (m->dtype == NKA_F32)
    st = ffnn_eloc_impl<float>(*m, op_kind, op, sigma, n, eloc, (float*)log_psi, (unsigned char*)workspace,
                               workspace_bytes, s);
  else
    st = ffnn_eloc_impl<double>(*m, op_kind, op, sigma, n, eloc, (double*)log_psi, (unsigned char*)workspace,
                                workspace_bytes, s);
  if (st) return st;
  NKA_CHECK_LAUNCH("nka_ffnn_local_energy");
  return NKA_OK;
}

int nka_ffnn_metropolis(const nka_ffnn* m, int rule, const int32_t* clusters, int n_clusters,
                        const uint32_t* cluster_weights, int8_t* sigma,
                        void* log_prob, int n_chains, int sweep_size, int n_discard, int chain_length,
                        double machine_pow, uint64_t seed, uint64_t counter, uint64_t* counter_dev,
                        int8_t* samples, void* log_prob_samples, int64_t* n_accepted, void* stream) {
  if (int st = check_model(m, "nka_ffnn_metropolis")) return st;
  NKA_CHECK_ARG(rule == NKA_RULE_LOCAL || rule == NKA_RULE_EXCHANGE, "nka_ffnn_metropolis: unknown rule %d",
                rule);
  NKA_CHECK_ARG(rule == NKA_RULE_LOCAL || (clusters && n_clusters > 0 && n_clusters <= FF_MAXCL * 64),
                "nka_ffnn_metropolis: exchange rule needs 1..%d clusters", FF_MAXCL * 64);
  NKA_CHECK_ARG(n_chains >= 0 && sweep_size >= 0 && n_discard >= 0 && chain_length >= 0,
                "nka_ffnn_metropolis: bad sizes");
  NKA_CHECK_ARG(n_chains == 0 || (sigma && log_prob && n_accepted && (chain_length == 0 || samples)),
                "nka_ffnn_metropolis: null buffer");
  if (n_chains == 0) return NKA_OK;
  FfSweepArgs sa;
  sa.rule = rule;
  sa.clusters = clusters;
  sa.n_clusters = rule == NKA_RULE_EXCHANGE ? n_clusters : 0;
  sa.weights = rule == NKA_RULE_EXCHANGE ? cluster_weights : nullptr;
  sa.C = n_chains;
  sa.sweep_size = sweep_size;
  sa.n_discard = n_discard;
  sa.chain_length = chain_length;
  sa.machine_pow = machine_pow;
  sa.k0 = (uint32_t)seed;
  sa.k1 = (uint32_t)(seed >> 32);
  sa.counter = counter;
  sa.counter_dev = counter_dev;
  const int ncl = (rule == NKA_RULE_EXCHANGE && !cluster_weights) ? n_clusters : 0;
  const int64_t lds = m->dtype == NKA_F32 ? (int64_t)metropolis_lds<float>(*m, ncl)
                                          : (int64_t)metropolis_lds<double>(*m, ncl);
  NKA_CHECK_ARG(lds <= FF_LDS_MAX, "nka_ffnn_metropolis: %d clusters need %lld bytes of LDS (max %lld)",
                n_clusters, (long long)lds, (long long)FF_LDS_MAX);
  hipStream_t s = (hipStream_t)stream;
  if (m->dtype == NKA_F32)
    dispatch_ffnn_sweep<float>(*m, sa, sigma, log_prob, samples, log_prob_samples, n_accepted, s);
  else
    dispatch_ffnn_sweep<double>(*m, sa, sigma, log_prob, samples, log_prob_samples, n_accepted, s);
  NKA_CHECK_LAUNCH("nka_ffnn_metropolis");
  return NKA_OK;
}

int64_t nka_ffnn_prepare_workspace(const nka_ffnn* m, int64_t ns) {
  if (!m || ns <= 0) return 0;
  const int64_t ts = m->dtype == NKA_F32 ? 4 : 8;
  return ffpad(ns * m->n_hidden1 * ts) + ffpad(ns * ts) + 64;
}

int nka_ffnn_prepare(const nka_ffnn* m, const int8_t* sigma, int64_t ns, void* h1, void* d1, void* g2,
                     void* workspace, void* stream) {
  if (int st = check_model(m, "nka_ffnn_prepare")) return st;
  NKA_CHECK_ARG(ns >= 0 && (ns == 0 || (sigma && h1 && d1 && g2 && workspace)), "nka_ffnn_prepare: bad arguments");
  if (ns == 0) return NKA_OK;
  hipStream_t s = (hipStream_t)stream;
  const int64_t ts = m->dtype == NKA_F32 ? 4 : 8;
  unsigned char* w = (unsigned char*)workspace;
  void* th = w;
  void* lp = w + ffpad(ns * m->n_hidden1 * ts);
  const unsigned g = (unsigned)((ns + FF_ROWS - 1) / FF_ROWS);
  if (m->dtype == NKA_F32) {
    launch_forward0<float>(*m, sigma, ns, (float*)lp, (float*)th, (float*)g2, s);
    allow_lds(ffnn_backprop_kernel<float>, (int64_t)backprop_lds<float>(*m));
    hipLaunchKernelGGL((ffnn_backprop_kernel<float>), dim3(g), dim3(FF_THREADS), backprop_lds<float>(*m), s, *m,
                       (const float*)th, (const float*)g2, ns, (float*)h1, (float*)d1);
  } else {
    launch_forward0<double>(*m, sigma, ns, (double*)lp, (double*)th, (double*)g2, s);
    allow_lds(ffnn_backprop_kernel<double>, (int64_t)backprop_lds<double>(*m));
    hipLaunchKernelGGL((ffnn_backprop_kernel<double>), dim3(g), dim3(FF_THREADS), backprop_lds<double>(*m), s, *m,
                       (const double*)th, (const double*)g2, ns, (double*)h1, (double*)d1);
  }
  NKA_CHECK_LAUNCH("nka_ffnn_prepare");
  return NKA_OK;
}

int64_t nka_ffnn_contract_workspace(const nka_ffnn* m, int64_t ns, int n_w) {
  if (!m || ns <= 0) return 0;
  const int64_t nch = (ns + CT_S - 1) / CT_S;
  const int64_t biggest = std::max<int64_t>((int64_t)m->n_visible * m->n_hidden1,
                                            (int64_t)m->n_hidden1 * m->n_hidden2);
  return nch * biggest * 8 + 64 + (int64_t)n_w * 0;
}

int nka_ffnn_contract(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1, const void* g2,
                      int64_t ns, const double* w, int n_w, const double* mean, double* out, void* workspace,
                      void* stream) {
  if (int st = check_model(m, "nka_ffnn_contract")) return st;
  NKA_CHECK_ARG(ns > 0 && n_w >= 1 && n_w <= 2 && sigma && h1 && d1 && g2 && w && out && workspace,
                "nka_ffnn_contract: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int64_t np = ffnn_np(*m);
  const int64_t nch = (ns + CT_S - 1) / CT_S;
  const int64_t biggest = std::max<int64_t>((int64_t)m->n_visible * m->n_hidden1,
                                            (int64_t)m->n_hidden1 * m->n_hidden2);
  double* part = (double*)workspace;
  double* wsum = part + nch * biggest;
  for (int vi = 0; vi < n_w; ++vi) {
    if (m->dtype == NKA_F32)
      contract_one<float>(*m, sigma, (const float*)h1, (const float*)d1, (const float*)g2, ns, w + vi * ns, mean,
                          out + vi * np, part, wsum, s);
    else
      contract_one<double>(*m, sigma, (const double*)h1, (const double*)d1, (const double*)g2, ns, w + vi * ns,
                           mean, out + vi * np, part, wsum, s);
  }
  NKA_CHECK_LAUNCH("nka_ffnn_contract");
  return NKA_OK;
}

static int64_t ffnn_ntk_base_bytes(const nka_ffnn& m, int64_t ns) {
  const int64_t H1 = m.n_hidden1, H2 = m.n_hidden2, N = m.n_visible;
  return (ns * (H1 + H2 + H1) + ns * N + ns * H1 + ns + 8) * 8;
}
static int64_t ffnn_ld8(int h) { return (h + 7) / 8 * 8; }

int64_t nka_ffnn_ntk_workspace(const nka_ffnn* m, int64_t ns) {
  if (!m || ns <= 0) return 0;
  int64_t b = ffnn_ntk_base_bytes(*m, ns) + 256;
  if (m->dtype == NKA_F32) {
    // bf16 hi + lo planes of h1, g2 and d1, and the packed spins
    const int64_t ld = 2 * ffnn_ld8(m->n_hidden1) + ffnn_ld8(m->n_hidden2);
    b += (2 * ns * ld * (int64_t)sizeof(uint16_t) + 255) / 256 * 256 + 256;
    b += ns * ((m->n_visible + 31) / 32) * (int64_t)sizeof(uint32_t);
  }
  return b;
}

int nka_ffnn_ntk(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1, const void* g2, int64_t ns,
                 const double* mean, double scale, double shift, double offset, double* out, int64_t ldo,
                 void* workspace, void* stream) {
  if (int st = check_model(m, "nka_ffnn_ntk")) return st;
  NKA_CHECK_ARG(ns > 0 && sigma && h1 && d1 && g2 && mean && out && workspace && ldo >= ns,
                "nka_ffnn_ntk: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int N = m->n_visible, H1 = m->n_hidden1, H2 = m->n_hidden2;
  double* h1d = (double*)workspace;
  double* g2d = h1d + ns * H1;
  double* d1d = g2d + ns * H2;
  double* U1 = d1d + ns * H1;
  double* U2 = U1 + ns * N;
  double* vv = U2 + ns * H1;
  double* cc = vv + ns;
  const unsigned gc = 4096;
  if (m->dtype == NKA_F32) {
    hipLaunchKernelGGL((ffnn_to_f64_kernel<float>), dim3(gc), dim3(256), 0, s, (const float*)h1, ns * H1, h1d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<float>), dim3(gc), dim3(256), 0, s, (const float*)g2, ns * H2, g2d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<float>), dim3(gc), dim3(256), 0, s, (const float*)d1, ns * H1, d1d);
  } else {
    hipLaunchKernelGGL((ffnn_to_f64_kernel<double>), dim3(gc), dim3(256), 0, s, (const double*)h1, ns * H1, h1d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<double>), dim3(gc), dim3(256), 0, s, (const double*)g2, ns * H2, g2d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<double>), dim3(gc), dim3(256), 0, s, (const double*)d1, ns * H1, d1d);
  }
  // U1 = D1 M_W1^T (ns x N), U2 = G2 M_W2^T (ns x H1)
  int st = nka_gemm_nt_f64(d1d, H1, mean + m->w1_off, H1, U1, N, ns, N, H1, 1.0, 0.0, stream);
  if (st) return st;
  st = nka_gemm_nt_f64(g2d, H2, mean + m->w2_off, H2, U2, H1, ns, H1, H2, 1.0, 0.0, stream);
  if (st) return st;
  const int64_t np = ffnn_np(*m);
  hipLaunchKernelGGL(ffnn_ntk_vec_kernel, dim3((unsigned)std::min<int64_t>((ns + 3) / 4, 4096)), dim3(256), 0, s,
                     h1d, g2d, sigma, d1d, U1, U2, N, H1, H2, ns, mean, m->b1_off, m->b2_off, np, vv, cc);
  const double hb1 = m->b1_off >= 0 ? 1.0 : 0.0, hb2 = m->b2_off >= 0 ? 1.0 : 0.0;
  if (m->dtype == NKA_F32) {
    // H, G and D on bf16 matrix cores (three products per operand pair, FP32
    // accumulation), X and the centring exact in the epilogue
    unsigned char* p = (unsigned char*)workspace + (ffnn_ntk_base_bytes(*m, ns) + 255) / 256 * 256;
    const int64_t l1 = ffnn_ld8(H1), l2 = ffnn_ld8(H2);
    uint16_t* h1h = (uint16_t*)p;
    uint16_t* h1l = h1h + ns * l1;
    uint16_t* g2h = h1l + ns * l1;
    uint16_t* g2l = g2h + ns * l2;
    uint16_t* d1h = g2l + ns * l2;
    uint16_t* d1l = d1h + ns * l1;
    uint32_t* bits = (uint32_t*)(p + (2 * ns * (2 * l1 + l2) * (int64_t)sizeof(uint16_t) + 255) / 256 * 256);
    const int nw = (N + 31) / 32;
    launch_ntk_split((const float*)h1, ns, H1, (int)l1, h1h, h1l, s);
    launch_ntk_split((const float*)g2, ns, H2, (int)l2, g2h, g2l, s);
    launch_ntk_split((const float*)d1, ns, H1, (int)l1, d1h, d1l, s);
    launch_ntk_pack_spins(sigma, ns, N, nw, bits, s);
    const int64_t T2 = (ns + 127) / 128;
    NKA_CHECK_ARG(T2 * (T2 + 1) / 2 < (1LL << 31), "nka_ffnn_ntk: too large");
    if (launch_gram_ffnn(h1h, h1l, l1, l1, g2h, g2l, l2, l2, d1h, d1l, l1, l1, ns, bits, nw, N, vv, cc, hb1, hb2,
                         scale, shift, offset, out, ldo, s)) {
      NKA_CHECK_LAUNCH("nka_ffnn_ntk");
      return NKA_OK;
    }
  }
  const int64_t T = (ns + NF_T - 1) / NF_T;
  const int64_t tiles = T * (T + 1) / 2;
  NKA_CHECK_ARG(tiles < (1LL << 31), "nka_ffnn_ntk: too large");
  const unsigned grid = (unsigned)(((tiles + 7) / 8) * 8);
  hipLaunchKernelGGL(ffnn_ntk_kernel, dim3(grid), dim3(256), 0, s, h1d, g2d, sigma, d1d, N, H1, H2, ns, vv, cc,
                     m->b1_off >= 0 ? 1.0 : 0.0, m->b2_off >= 0 ? 1.0 : 0.0, scale, shift, offset, out, ldo, tiles);
  NKA_CHECK_LAUNCH("nka_ffnn_ntk");
  return NKA_OK;
}

int64_t nka_ffnn_jvp_workspace(const nka_ffnn* m, int64_t ns) { return nka_ffnn_ntk_workspace(m, ns); }

// (O v)_s = d1_s.(v_b1 + x_s v_W1) + g2_s.(v_b2 + h1_s v_W2): the products
// D1 v_W1^T and G2 v_W2^T on FP64 matrix cores and the per-sample sums of the
// NTK's v = O mu with v in place of mu.
int nka_ffnn_jvp(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1, const void* g2,
                 int64_t ns, const double* v, const double* mean, double alpha, double beta, double* y,
                 void* workspace, void* stream) {
  if (int st = check_model(m, "nka_ffnn_jvp")) return st;
  NKA_CHECK_ARG(ns > 0 && sigma && h1 && d1 && g2 && v && y && workspace, "nka_ffnn_jvp: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int N = m->n_visible, H1 = m->n_hidden1, H2 = m->n_hidden2;
  double* h1d = (double*)workspace;
  double* g2d = h1d + ns * H1;
  double* d1d = g2d + ns * H2;
  double* U1 = d1d + ns * H1;
  double* U2 = U1 + ns * N;
  double* raw = U2 + ns * H1;
  double* c = raw + ns;
  const unsigned gc = 4096;
  if (m->dtype == NKA_F32) {
    hipLaunchKernelGGL((ffnn_to_f64_kernel<float>), dim3(gc), dim3(256), 0, s, (const float*)h1, ns * H1, h1d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<float>), dim3(gc), dim3(256), 0, s, (const float*)g2, ns * H2, g2d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<float>), dim3(gc), dim3(256), 0, s, (const float*)d1, ns * H1, d1d);
  } else {
    hipLaunchKernelGGL((ffnn_to_f64_kernel<double>), dim3(gc), dim3(256), 0, s, (const double*)h1, ns * H1, h1d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<double>), dim3(gc), dim3(256), 0, s, (const double*)g2, ns * H2, g2d);
    hipLaunchKernelGGL((ffnn_to_f64_kernel<double>), dim3(gc), dim3(256), 0, s, (const double*)d1, ns * H1, d1d);
  }
  int st = nka_gemm_nt_f64(d1d, H1, v + m->w1_off, H1, U1, N, ns, N, H1, 1.0, 0.0, stream);
  if (st) return st;
  st = nka_gemm_nt_f64(g2d, H2, v + m->w2_off, H2, U2, H1, ns, H1, H2, 1.0, 0.0, stream);
  if (st) return st;
  const int64_t np = ffnn_np(*m);
  hipLaunchKernelGGL(ffnn_ntk_vec_kernel, dim3((unsigned)std::min<int64_t>((ns + 3) / 4, 4096)), dim3(256), 0, s,
                     h1d, g2d, sigma, d1d, U1, U2, N, H1, H2, ns, v, m->b1_off, m->b2_off, np, raw, c);
  launch_jvp_finish(raw, ns, mean, v, np, c + 1, alpha, beta, y, s);
  NKA_CHECK_LAUNCH("nka_ffnn_jvp");
  return NKA_OK;
}

int nka_ffnn_jacobian(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1, const void* g2,
                      int64_t ns, const double* mean, double scale, int layout, int out_kind, void* out_a,
                      void* out_b, int64_t ldo, void* stream) {
  if (int st = check_model(m, "nka_ffnn_jacobian")) return st;
  NKA_CHECK_ARG(ns > 0 && sigma && h1 && d1 && g2 && out_a, "nka_ffnn_jacobian: bad arguments");
  NKA_CHECK_ARG(layout == NKA_LAYOUT_PS || layout == NKA_LAYOUT_SP, "nka_ffnn_jacobian: bad layout");
  NKA_CHECK_ARG(out_kind == NKA_OUT_F64 || out_kind == NKA_OUT_F32 || (out_kind == NKA_OUT_BF16X2 && out_b),
                "nka_ffnn_jacobian: bad output kind");
  const int64_t np = ffnn_np(*m);
  NKA_CHECK_ARG(ldo >= (layout == NKA_LAYOUT_PS ? ns : np), "nka_ffnn_jacobian: ldo too small");
  hipStream_t s = (hipStream_t)stream;
  const int64_t total = ns * np;
  const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 1 << 16);
  if (m->dtype == NKA_F32)
    hipLaunchKernelGGL((ffnn_jacobian_kernel<float>), dim3(g), dim3(256), 0, s, *m, sigma, (const float*)h1,
                       (const float*)d1, (const float*)g2, ns, np, mean, scale, layout, out_kind, out_a, out_b, ldo);
  else
    hipLaunchKernelGGL((ffnn_jacobian_kernel<double>), dim3(g), dim3(256), 0, s, *m, sigma, (const double*)h1,
                       (const double*)d1, (const double*)g2, ns, np, mean, scale, layout, out_kind, out_a, out_b,
                       ldo);
  NKA_CHECK_LAUNCH("nka_ffnn_jacobian");
  return NKA_OK;
}

}  // extern "C"
