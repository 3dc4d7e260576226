// RBM forward, fused local energy and Jacobian kernels.
//
// nka_rbm_log_psi       RBM.__call__ (netket/models/rbm.py:57-81)
// nka_rbm_local_energy  local_value_kernel_jax + get_conn_padded
//                       (netket/vqs/mc/kernels.py:62-71)
// nka_rbm_prepare       tanh(theta) used by the analytic Jacobian
// nka_rbm_contract      sum_s O[s,p] w[v,s]: centring means and forces
//                       (netket/jax/_jacobian/logic.py:365-369,
//                        netket/vqs/mc/mc_state/expect_forces.py:98-104)
// nka_rbm_jacobian      centred, rescaled Jacobian in either layout
//                       (netket/jax/_jacobian/logic.py:356-373)
#include "nka_rbm.hpp"

namespace nka {

// ---------------------------------------------------------------------------
// log psi
// ---------------------------------------------------------------------------
template <class T, int JL, bool STAGE>
__global__ void __launch_bounds__(256) rbm_log_psi_kernel(nka_rbm m, const int8_t* __restrict__ sigma,
                                                          int64_t n, T* __restrict__ out) {
  extern __shared__ unsigned char smem[];
  const RbmView<T> v = rbm_view<T>(m);
  const int wid = threadIdx.x / NKA_WAVE, waves = blockDim.x / NKA_WAVE;
  const size_t wbytes = STAGE ? pad16((size_t)v.N * v.M * sizeof(T)) : 0;
  int8_t* xs = (int8_t*)(smem + wbytes) + wid * pad16(v.N);
  const T* W = v.W;
  if (STAGE) {
    stage_w<T>((T*)smem, v.W, v.N, v.M);
    W = (const T*)smem;
  }
  for (int64_t b = (int64_t)blockIdx.x * waves + wid; b < n; b += (int64_t)gridDim.x * waves) {
    for (int i = lane_id(); i < v.N; i += NKA_WAVE) xs[i] = sigma[b * v.N + i];
    wave_lds_sync();
    HiddenState<T, JL> h;
    compute_theta<T, JL>(h, xs, W, v.b, v.N, v.M);
    const T lp = sum_log_cosh<T, JL>(h) + visible_term<T>(xs, v.a, v.N);
    if (lane_id() == 0) out[b] = lp;
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Local energy, one wavefront per sample.
// ---------------------------------------------------------------------------
struct OpArgs {
  int kind;
  nka_ising ising;
  nka_localop local;
};

__device__ inline double load_mel(const void* p, int64_t i, int dtype) {
  return dtype == NKA_F64 ? ((const double*)p)[i] : (double)((const float*)p)[i];
}

template <class T, int JL, bool STAGE>
__device__ __forceinline__ void rbm_eloc_body(const nka_rbm& m, const OpArgs& op, const int8_t* __restrict__ sigma,
                                              int64_t n, double* __restrict__ eloc, T* __restrict__ log_psi,
                                              int bx, int gx) {
  extern __shared__ unsigned char smem[];
  const RbmView<T> v = rbm_view<T>(m);
  const int N = v.N, M = v.M;
  const int wid = threadIdx.x / NKA_WAVE, waves = blockDim.x / NKA_WAVE;
  const int lane = lane_id();
  const size_t wbytes = STAGE ? pad16((size_t)N * M * sizeof(T)) : 0;
  int8_t* xs = (int8_t*)(smem + wbytes) + wid * pad16(N);
  const T* W = v.W;
  if (STAGE) {
    stage_w<T>((T*)smem, v.W, N, M);
    W = (const T*)smem;
  }
  for (int64_t b = (int64_t)bx * waves + wid; b < n; b += (int64_t)gx * waves) {
    for (int i = lane; i < N; i += NKA_WAVE) xs[i] = sigma[b * N + i];
    wave_lds_sync();
    HiddenState<T, JL> h;
    compute_theta<T, JL>(h, xs, W, v.b, N, M);
    if (log_psi) {
      const T lp = sum_log_cosh<T, JL>(h) + visible_term<T>(xs, v.a, N);
      if (lane == 0) log_psi[b] = lp;
    }
    T nth[JL], nlc[JL];
    double E = 0.0;
    if (op.kind == NKA_OP_ISING) {
      const nka_ising& is = op.ising;
      int zz = 0;
      for (int e = lane; e < is.n_edges; e += NKA_WAVE)
        zz += (xs[is.edges[2 * e]] == xs[is.edges[2 * e + 1]]) ? 1 : -1;
      zz = wave_sum(zz);
      // mel of row 0 is J * (integer sum) in the operator dtype
      const double diag =
          is.mel_dtype == NKA_F64 ? is.J * (double)zz : (double)((float)is.J * (float)zz);
      E += diag;
      if (is.h != 0.0) {
        const double mh = is.mel_dtype == NKA_F64 ? -is.h : (double)(-(float)is.h);
        for (int r = 0; r < N; ++r) {
          const int sites[1] = {r};
          const T d = flip_delta<T, JL>(h, xs, W, v.a, N, M, sites, 1, nth, nlc);
          E += mh * (double)exp_t<T>(d);
        }
      }
    } else {
      const nka_localop& lo = op.local;
      // diagonal: constant + sum over terms of the selected row's diagonal element
      double dpart = 0.0;
      for (int t = lane; t < lo.n_ops; t += NKA_WAVE) {
        const int ns = lo.op_nsites[t];
        int r = 0;
        for (int k = 0; k < ns; ++k) r = 2 * r + (xs[lo.op_sites[4 * t + k]] < 0 ? 1 : 0);
        dpart += load_mel(lo.row_diag, lo.op_row_off[t] + r, lo.mel_dtype);
      }
      const double diag = lo.constant + wave_sum(dpart);
      if (lo.nonzero_diagonal && fabs(diag) > lo.mel_cutoff) E += diag;
      for (int t = 0; t < lo.n_ops; ++t) {
        const int ns = lo.op_nsites[t];
        int st[4];
        int r = 0;
        for (int k = 0; k < ns; ++k) {
          st[k] = lo.op_sites[4 * t + k];
          r = 2 * r + (xs[st[k]] < 0 ? 1 : 0);
        }
        r = uniform_value<int>(r);
        const int row = lo.op_row_off[t] + r;
        const int nc = lo.row_nconn[row];
        const int off = lo.row_conn_off[row];
        for (int c = 0; c < nc; ++c) {
          const int nidx = lo.conn_newidx[off + c];
          int fl[4];
          int nf = 0;
          for (int k = 0; k < ns; ++k)
            if ((((nidx ^ r) >> (ns - 1 - k)) & 1) != 0) fl[nf++] = st[k];
          const T d = flip_delta<T, JL>(h, xs, W, v.a, N, M, fl, nf, nth, nlc);
          E += load_mel(lo.conn_mel, off + c, lo.mel_dtype) * (double)exp_t<T>(d);
        }
      }
    }
    if (lane == 0) eloc[b] = E;
    wave_lds_sync();
  }
}

template <class T, int JL, bool STAGE>
__global__ void __launch_bounds__(256) rbm_eloc_kernel(nka_rbm m, OpArgs op, const int8_t* __restrict__ sigma,
                                                       int64_t n, double* __restrict__ eloc,
                                                       T* __restrict__ log_psi) {
  rbm_eloc_body<T, JL, STAGE>(m, op, sigma, n, eloc, log_psi, blockIdx.x, gridDim.x);
}

// The local energies of a batch of jobs of one shape (Ising operators on the
// same lattice, each with its own couplings): blockIdx.y selects the job.
template <class T, int JL, bool STAGE>
__global__ void __launch_bounds__(256) rbm_eloc_kernel_batched(const nka_eloc_job* __restrict__ jobs, int64_t n) {
  const nka_eloc_job& j = jobs[blockIdx.y];
  OpArgs op{};
  op.kind = NKA_OP_ISING;
  op.ising = j.op;
  rbm_eloc_body<T, JL, STAGE>(j.model, op, j.sigma, n, j.eloc, (T*)j.log_psi, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// t = tanh(theta): row-major (ns, M) and transposed (M, ns).
// ---------------------------------------------------------------------------
template <class T, int JL, bool STAGE>
__global__ void __launch_bounds__(256) rbm_prepare_kernel(nka_rbm m, const int8_t* __restrict__ sigma, int64_t ns,
                                                          T* __restrict__ t, T* __restrict__ tT) {
  extern __shared__ unsigned char smem[];
  const RbmView<T> v = rbm_view<T>(m);
  const int N = v.N, M = v.M;
  const int wid = threadIdx.x / NKA_WAVE, waves = blockDim.x / NKA_WAVE;
  const int lane = lane_id();
  const size_t wbytes = STAGE ? pad16((size_t)N * M * sizeof(T)) : 0;
  int8_t* xs = (int8_t*)(smem + wbytes) + wid * pad16(N);
  const T* W = v.W;
  if (STAGE) {
    stage_w<T>((T*)smem, v.W, N, M);
    W = (const T*)smem;
  }
  for (int64_t b = (int64_t)blockIdx.x * waves + wid; b < ns; b += (int64_t)gridDim.x * waves) {
    for (int i = lane; i < N; i += NKA_WAVE) xs[i] = sigma[b * N + i];
    wave_lds_sync();
    HiddenState<T, JL> h;
    compute_theta<T, JL>(h, xs, W, v.b, N, M);
#pragma unroll
    for (int jj = 0; jj < JL; ++jj) {
      const int j = lane + NKA_WAVE * jj;
      if (j < M) {
        const T tv = tanh_t<T>(h.th[jj]);
        if (t) t[b * M + j] = tv;
        if (tT) tT[(int64_t)j * ns + b] = tv;
      }
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Weighted contraction over samples, split over sample chunks:
//   part[c, v, p] = sum_{s in chunk c} O[s,p] w[v,s]
// O[s, b_j] = t_sj, O[s, W_ij] = s_i t_sj, O[s, a_i] = s_i.
// A workgroup owns a chunk of nsub x 64 samples and a tile of 64 hidden
// units.  For each 64-row slab of visible units it walks the chunk in 64-sample
// steps: the step's weights, spins and tanh values are staged in LDS (every
// global load of the stage in flight at once), then the four waves form the
// W entries of the slab on FP64 matrix cores (K = samples).
// nsub > 1 for models with many hidden units, so that the partials written and
// re-read by the reduction stay a small multiple of the parameter count.
// ---------------------------------------------------------------------------
typedef double ddx4 __attribute__((ext_vector_type(4)));
constexpr int kContractChunk = 64;    // samples per staged step
constexpr int kContractMaxW = 2;      // weight vectors per call

// 64-sample steps per chunk: about 2048 workgroups over chunks x hidden tiles
static int contract_nsub(const nka_rbm& m, int64_t ns) {
  const int64_t steps = (ns + kContractChunk - 1) / kContractChunk;
  const int64_t jt = (m.n_hidden + 63) / 64;
  return (int)std::max<int64_t>(1, std::min<int64_t>(steps, steps * jt / 2048));
}
static int64_t contract_chunks(const nka_rbm& m, int64_t ns) {
  const int64_t cs = (int64_t)kContractChunk * contract_nsub(m, ns);
  return (ns + cs - 1) / cs;
}

template <class T, int NW>
__global__ void __launch_bounds__(256) rbm_contract_partial(nka_rbm m, const int8_t* __restrict__ sigma,
                                                            const T* __restrict__ t, int64_t ns,
                                                            const double* __restrict__ w,
                                                            double* __restrict__ part, int64_t np, int nsub) {
  __shared__ double ws[NW][kContractChunk];
  __shared__ int8_t sg[kContractChunk][64];
  __shared__ T tl[kContractChunk][64];
  const int N = m.n_visible, M = m.n_hidden;
  const int64_t c0 = (int64_t)blockIdx.x * kContractChunk * nsub;
  const int lane = lane_id();
  const int j = blockIdx.y * 64 + lane;
  const int wid = threadIdx.x / 64;
  // partial layout: [chunk][v][np + 1]; slot np holds the chunk's sum of w[v]
  const int64_t ld = np + 1;
  double* out = part + (int64_t)blockIdx.x * NW * ld;
  double b2[NW];
#pragma unroll
  for (int vv = 0; vv < NW; ++vv) b2[vv] = 0.0;
  for (int i0 = 0; i0 < N; i0 += 64) {
    const int ni = min(64, N - i0);
    // W entries of the slab on FP64 MFMA: wave (wm, wn) owns a 32 x 32 block
    // of (visible row, hidden unit) as 2 x 2 tiles of 16 x 16;
    // A[i][s] = sigma_si w_s, B[s][j] = t_sj, K = samples
    ddx4 acc[NW][2][2];
    double a2[NW];
#pragma unroll
    for (int vv = 0; vv < NW; ++vv) {
      a2[vv] = 0.0;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[vv][a][b] = (ddx4){0.0, 0.0, 0.0, 0.0};
    }
    const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
    for (int sb = 0; sb < nsub; ++sb) {
      const int64_t s0 = c0 + (int64_t)sb * kContractChunk;
      if (s0 >= ns) break;
      const int cn = (int)min((int64_t)kContractChunk, ns - s0);
      __syncthreads();
      {
        // wave w loads samples w + 4q (q < 16): tanh values along j, spins along i
        T tv[16];
        int8_t sv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int s = wid + 4 * q;
          tv[q] = (s < cn && j < M) ? t[(s0 + s) * M + j] : (T)0;
          sv[q] = (s < cn && lane < ni) ? sigma[(s0 + s) * N + i0 + lane] : (int8_t)0;
        }
        if (threadIdx.x < NW * kContractChunk) {
          const int vv = threadIdx.x / kContractChunk, s = threadIdx.x % kContractChunk;
          ws[vv][s] = s < cn ? w[(int64_t)vv * ns + s0 + s] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          tl[wid + 4 * q][lane] = tv[q];
          sg[wid + 4 * q][lane] = sv[q];
        }
      }
      __syncthreads();
#pragma unroll 4
      for (int ks = 0; ks < kContractChunk; ks += 4) {
        // operand element of lane l: A[row l%16][k l/16], B[k l/16][col l%16]
        const int sk = ks + (lane >> 4);
        double bv[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = (double)tl[sk][wn + b * 16 + (lane & 15)];
        int8_t sv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) sv[a] = sg[sk][wm + a * 16 + (lane & 15)];
#pragma unroll
        for (int vv = 0; vv < NW; ++vv) {
          const double wv = ws[vv][sk];
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const double av = sv[a] > 0 ? wv : (sv[a] < 0 ? -wv : 0.0);
#pragma unroll
            for (int b = 0; b < 2; ++b)
              acc[vv][a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[b], acc[vv][a][b], 0, 0, 0);
          }
        }
      }
      // visible-bias entries of this slab (tile 0), hidden-bias entries (first slab)
      if (m.visible_off >= 0 && blockIdx.y == 0 && threadIdx.x < ni) {
        const int ii = threadIdx.x;
        for (int s = 0; s < cn; ++s) {
          const double sv = sg[s][ii] > 0 ? 1.0 : -1.0;
#pragma unroll
          for (int vv = 0; vv < NW; ++vv) a2[vv] += sv * ws[vv][s];
        }
      }
      if (i0 == 0 && m.bias_off >= 0 && wid == 0 && j < M) {
        for (int s = 0; s < cn; ++s) {
          const double tvs = (double)tl[s][lane];
#pragma unroll
          for (int vv = 0; vv < NW; ++vv) b2[vv] += tvs * ws[vv][s];
        }
      }
    }
    // element e of lane l in a 16x16 tile is (row l/16 + 4e, col l%16)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int jj = blockIdx.y * 64 + wn + b * 16 + (lane & 15);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int ii = wm + a * 16 + (lane >> 4) + 4 * e;
          if (ii < ni && jj < M) {
            const int64_t p = m.kernel_off + (int64_t)(i0 + ii) * M + jj;
#pragma unroll
            for (int vv = 0; vv < NW; ++vv) out[(int64_t)vv * ld + p] = acc[vv][a][b][e];
          }
        }
      }
    if (m.visible_off >= 0 && blockIdx.y == 0 && threadIdx.x < ni) {
#pragma unroll
      for (int vv = 0; vv < NW; ++vv) out[(int64_t)vv * ld + m.visible_off + i0 + threadIdx.x] = a2[vv];
    }
  }
  // per-chunk weight sums (used to centre the contraction)
  if (blockIdx.y == 0 && wid == 1) {
    const int64_t cend = min(ns, c0 + (int64_t)kContractChunk * nsub);
#pragma unroll
    for (int vv = 0; vv < NW; ++vv) {
      double sw = 0.0;
      for (int64_t s = c0 + lane; s < cend; s += 64) sw += w[(int64_t)vv * ns + s];
      sw = wave_sum(sw);
      if (lane == 0) out[(int64_t)vv * ld + np] = sw;
    }
  }
  if (m.bias_off >= 0 && wid == 0 && j < M) {
#pragma unroll
    for (int vv = 0; vv < NW; ++vv) out[(int64_t)vv * ld + m.bias_off + j] = b2[vv];
  }
}

__global__ void __launch_bounds__(256) contract_reduce(const double* __restrict__ part, int n_chunks, int n_w,
                                                       int64_t np, const double* __restrict__ mean,
                                                       double* __restrict__ out) {
  // out[v, p] = sum_c part[c, v, p] - mean[p] * sum_c part[c, v, np]
  const int64_t ld = np + 1;
  const int64_t len = (int64_t)n_w * np;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < len;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = q / np, p = q % np;
    double acc = 0.0, sw = 0.0;
    for (int c = 0; c < n_chunks; ++c) {
      acc += part[((int64_t)c * n_w + v) * ld + p];
      if (mean) sw += part[((int64_t)c * n_w + v) * ld + np];
    }
    out[q] = mean ? acc - mean[p] * sw : acc;
  }
}

// ---------------------------------------------------------------------------
// Centred Jacobian writer.
// ---------------------------------------------------------------------------
__device__ inline void store_out(int kind, void* a, void* b, int64_t idx, double val) {
  if (kind == NKA_OUT_F64) {
    ((double*)a)[idx] = val;
  } else if (kind == NKA_OUT_F32) {
    ((float*)a)[idx] = (float)val;
  } else {
    const float f = (float)val;
    const uint16_t hi = bf16_rne(f);
    const uint16_t lo = bf16_rne(f - bf16_to_f(hi));
    ((uint16_t*)a)[idx] = hi;
    ((uint16_t*)b)[idx] = lo;
  }
}

// Parameter-major layout: X[p, s].  A workgroup writes the rows of one
// parameter group for a chunk of 512 samples: group g < N is the W block of
// visible unit i = g (rows kernel_off + i M + j, j < M), group N the hidden
// bias (rows bias_off + j), group N + 1 the visible bias (rows visible_off +
// i).  Thread t covers 8 consecutive samples (s0 + 8 (t % 64) ..) of the rows
// r = t / 64 + 4k, so each row is written as 8-element runs (one 16-byte
// store per bf16 plane) and tanh values are read as 8-element runs of the
// transposed tanh tT[j, s]; the chunk's spins of unit i are staged in LDS.
// VEC: ns is a multiple of 8 (aligned 8-element runs of tT).
constexpr int kJacChunk = 512;

template <class T, bool VEC>
__global__ void __launch_bounds__(256) rbm_jac_ps_kernel(nka_rbm m, const int8_t* __restrict__ sigma,
                                                         const T* __restrict__ tT, int64_t ns,
                                                         const double* __restrict__ mean, double scale,
                                                         int kind, void* out_a, void* out_b, int64_t np,
                                                         int64_t ldo) {
  __shared__ int8_t sgi[kJacChunk];
  const int N = m.n_visible, M = m.n_hidden;
  const int g = blockIdx.y;
  const int64_t s0 = (int64_t)blockIdx.x * kJacChunk;
  const int tid = threadIdx.x;
  const int s8 = (tid & 63) * 8;
  const int rsub = tid >> 6;
  if (g < N) {
    for (int q = tid; q < kJacChunk; q += 256)
      sgi[q] = (s0 + q < ns) ? sigma[(s0 + q) * N + g] : (int8_t)1;
    __syncthreads();
  } else if ((g == N && m.bias_off < 0) || (g == N + 1 && m.visible_off < 0)) {
    return;
  }
  const int rows = (g == N + 1) ? N : M;
  const bool full = s0 + s8 + 8 <= ns;
  for (int r = rsub; r < rows; r += 4) {
    int64_t p;
    T tv[8];
    if (g == N + 1) {
      p = m.visible_off + r;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int64_t s = s0 + s8 + e;
        tv[e] = (s < ns && sigma[s * N + r] > 0) ? (T)1 : (T)-1;
      }
    } else {
      p = (g == N) ? m.bias_off + r : m.kernel_off + (int64_t)g * M + r;
      const T* src = tT + (int64_t)r * ns + s0 + s8;
      if (VEC && full) {
        if (sizeof(T) == 4) {
          const float4 a = *(const float4*)src, b = *(const float4*)(src + 4);
          tv[0] = a.x; tv[1] = a.y; tv[2] = a.z; tv[3] = a.w;
          tv[4] = b.x; tv[5] = b.y; tv[6] = b.z; tv[7] = b.w;
        } else {
#pragma unroll
          for (int e = 0; e < 8; e += 2) {
            const double2 a = *(const double2*)(src + e);
            tv[e] = (T)a.x;
            tv[e + 1] = (T)a.y;
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) tv[e] = (s0 + s8 + e < ns) ? src[e] : (T)0;
      }
      if (g < N) {
#pragma unroll
        for (int e = 0; e < 8; ++e) tv[e] = sgi[s8 + e] > 0 ? tv[e] : -tv[e];
      }
    }
    const double mp = mean[p];
    const int64_t o_idx = p * ldo + s0 + s8;
    if (kind == NKA_OUT_BF16X2 && full) {
      uint32_t hw[4], lw[4];
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        const float f0 = (float)(((double)tv[e] - mp) * scale);
        const float f1 = (float)(((double)tv[e + 1] - mp) * scale);
        const uint16_t h0 = bf16_rne(f0), h1 = bf16_rne(f1);
        const uint16_t l0 = bf16_rne(f0 - bf16_to_f(h0)), l1 = bf16_rne(f1 - bf16_to_f(h1));
        hw[e / 2] = (uint32_t)h0 | ((uint32_t)h1 << 16);
        lw[e / 2] = (uint32_t)l0 | ((uint32_t)l1 << 16);
      }
      *(uint4*)((uint16_t*)out_a + o_idx) = make_uint4(hw[0], hw[1], hw[2], hw[3]);
      *(uint4*)((uint16_t*)out_b + o_idx) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (s0 + s8 + e < ns) store_out(kind, out_a, out_b, o_idx + e, ((double)tv[e] - mp) * scale);
    }
  }
}

// Sample-major layout: X[s, p]; one thread per (s, p) with p fastest.
template <class T>
__global__ void __launch_bounds__(256) rbm_jac_sp_kernel(nka_rbm m, const int8_t* __restrict__ sigma,
                                                         const T* __restrict__ t, int64_t ns,
                                                         const double* __restrict__ mean, double scale,
                                                         int kind, void* out_a, void* out_b, int64_t np,
                                                         int64_t ldo) {
  const int N = m.n_visible, M = m.n_hidden;
  const int64_t total = np * ns;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx / np, p = idx % np;
    const int64_t o_idx = s * ldo + p;
    double o;
    if (m.bias_off >= 0 && p >= m.bias_off && p < m.bias_off + M) {
      o = (double)t[s * M + (p - m.bias_off)];
    } else if (p >= m.kernel_off && p < m.kernel_off + (int64_t)N * M) {
      const int64_t q = p - m.kernel_off;
      const int i = (int)(q / M), j = (int)(q % M);
      const double tv = (double)t[s * M + j];
      o = sigma[s * N + i] > 0 ? tv : -tv;
    } else {
      const int i = (int)(p - m.visible_off);
      o = sigma[s * N + i] > 0 ? 1.0 : -1.0;
    }
    store_out(kind, out_a, out_b, o_idx, (o - mean[p]) * scale);
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int64_t n_params(const nka_rbm& m) {
  return (m.bias_off >= 0 ? m.n_hidden : 0) + (int64_t)m.n_visible * m.n_hidden +
         (m.visible_off >= 0 ? m.n_visible : 0);
}

static int check_rbm(const nka_rbm* m, const char* who) {
  if (!m || !m->params) {
    set_error("%s: null model", who);
    return NKA_ERR_ARG;
  }
  if (m->n_visible <= 0 || m->n_hidden <= 0 || m->n_hidden > kMaxJL * NKA_WAVE) {
    set_error("%s: unsupported sizes N=%d M=%d (need 0 < M <= %d)", who, m->n_visible,
              m->n_hidden, kMaxJL * NKA_WAVE);
    return NKA_ERR_ARG;
  }
  if (m->dtype != NKA_F32 && m->dtype != NKA_F64) {
    set_error("%s: bad dtype", who);
    return NKA_ERR_ARG;
  }
  if (m->kernel_off < 0) {
    set_error("%s: kernel offset missing", who);
    return NKA_ERR_ARG;
  }
  return NKA_OK;
}

template <class T>
static bool stage_fits(const nka_rbm& m, int waves) {
  return pad16((size_t)m.n_visible * m.n_hidden * sizeof(T)) + waves * pad16(m.n_visible) <=
         kLdsBudget;
}

static unsigned grid_for(int64_t n, int waves, int64_t cap) {
  int64_t g = (n + waves - 1) / waves;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

template <int JL, class T>
static void launch_log_psi(const nka_rbm& m, const int8_t* sigma, int64_t n, T* out, hipStream_t s) {
  const int waves = 4;
  const bool stage = stage_fits<T>(m, waves);
  const size_t lds = (stage ? pad16((size_t)m.n_visible * m.n_hidden * sizeof(T)) : 0) +
                     waves * pad16(m.n_visible);
  const unsigned g = grid_for(n, waves, 2048);
  if (stage)
    hipLaunchKernelGGL((rbm_log_psi_kernel<T, JL, true>), dim3(g), dim3(64 * waves), lds, s, m,
                       sigma, n, out);
  else
    hipLaunchKernelGGL((rbm_log_psi_kernel<T, JL, false>), dim3(g), dim3(64 * waves), lds, s, m,
                       sigma, n, out);
}
template <int JL>
static void launch_log_psi_f32(const nka_rbm& m, const int8_t* sg, int64_t n, void* o, hipStream_t s) {
  launch_log_psi<JL, float>(m, sg, n, (float*)o, s);
}
template <int JL>
static void launch_log_psi_f64(const nka_rbm& m, const int8_t* sg, int64_t n, void* o, hipStream_t s) {
  launch_log_psi<JL, double>(m, sg, n, (double*)o, s);
}

template <int JL, class T>
static void launch_eloc(const nka_rbm& m, const OpArgs& op, const int8_t* sigma, int64_t n,
                        double* eloc, T* lp, hipStream_t s, const nka_eloc_job* jobs = nullptr,
                        int n_jobs = 0) {
  const int waves = 4;
  const bool stage = stage_fits<T>(m, waves);
  const size_t lds = (stage ? pad16((size_t)m.n_visible * m.n_hidden * sizeof(T)) : 0) +
                     waves * pad16(m.n_visible);
  const unsigned g = grid_for(n, waves, 2048);
  if (jobs) {
    // fewer blocks per job when the batch fills the device anyway
    const unsigned gb = std::max<unsigned>(1, std::min<unsigned>(g, (unsigned)((4096 + n_jobs - 1) / n_jobs)));
    if (stage)
      hipLaunchKernelGGL((rbm_eloc_kernel_batched<T, JL, true>), dim3(gb, n_jobs), dim3(64 * waves), lds, s,
                         jobs, n);
    else
      hipLaunchKernelGGL((rbm_eloc_kernel_batched<T, JL, false>), dim3(gb, n_jobs), dim3(64 * waves), lds, s,
                         jobs, n);
    return;
  }
  if (stage)
    hipLaunchKernelGGL((rbm_eloc_kernel<T, JL, true>), dim3(g), dim3(64 * waves), lds, s, m, op,
                       sigma, n, eloc, lp);
  else
    hipLaunchKernelGGL((rbm_eloc_kernel<T, JL, false>), dim3(g), dim3(64 * waves), lds, s, m, op,
                       sigma, n, eloc, lp);
}
template <int JL>
static void launch_eloc_f32(const nka_rbm& m, const OpArgs& op, const int8_t* sg, int64_t n,
                            double* e, void* lp, hipStream_t s, const nka_eloc_job* jobs, int n_jobs) {
  launch_eloc<JL, float>(m, op, sg, n, e, (float*)lp, s, jobs, n_jobs);
}
template <int JL>
static void launch_eloc_f64(const nka_rbm& m, const OpArgs& op, const int8_t* sg, int64_t n,
                            double* e, void* lp, hipStream_t s, const nka_eloc_job* jobs, int n_jobs) {
  launch_eloc<JL, double>(m, op, sg, n, e, (double*)lp, s, jobs, n_jobs);
}

// theta = sigma W + b for models whose W does not fit in LDS, as a tiled
// matrix product: 64 samples x 64 hidden units per workgroup, K = N in steps
// of 16 through LDS (sigma as exact +-1), so W is read once per 64 samples
// instead of once per sample.  FP32 on v_mfma_f32_32x32x2f32 (one 32x32
// block per wave), FP64 on v_mfma_f64_16x16x4f64 (2x2 blocks per wave).
// The epilogue writes t = tanh(theta) to t (ns, M) and/or tT (M, ns).
typedef float ffx16 __attribute__((ext_vector_type(16)));
constexpr int PG_T = 64, PG_K = 16;

template <class T, bool RAW>
__global__ void __launch_bounds__(256) rbm_prepare_gemm_kernel(nka_rbm m, const int8_t* __restrict__ sigma,
                                                               int64_t ns, T* __restrict__ t, T* __restrict__ tT) {
  __shared__ T sA[PG_T][PG_K + 1];
  __shared__ T sB[PG_K][PG_T + 1];
  const RbmView<T> v = rbm_view<T>(m);
  const int N = v.N, M = v.M;
  const int64_t s0 = (int64_t)blockIdx.x * PG_T;
  const int j0 = blockIdx.y * PG_T;
  const int tid = threadIdx.x, wid = tid >> 6, lane = lane_id();
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  ffx16 accf;
  ddx4 accd[2][2];
#pragma unroll
  for (int e = 0; e < 16; ++e) accf[e] = 0.0f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) accd[a][b] = (ddx4){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < N; k0 += PG_K) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + it * 256;
      const int r = idx >> 4, kc = idx & 15;            // A: 64 rows x 16 k
      const int kk = k0 + kc;
      T va = (T)0;
      if (s0 + r < ns && kk < N) va = sigma[(s0 + r) * N + kk] > 0 ? (T)1 : (T)-1;
      sA[r][kc] = va;
      const int kr = idx >> 6, c = idx & 63;            // B: 16 k x 64 cols
      T vb = (T)0;
      if (k0 + kr < N && j0 + c < M) vb = v.W[(int64_t)(k0 + kr) * M + j0 + c];
      sB[kr][c] = vb;
    }
    __syncthreads();
    if constexpr (sizeof(T) == 4) {
#pragma unroll
      for (int k = 0; k < PG_K; k += 2) {
        const float a = sA[wm + (lane & 31)][k + (lane >> 5)];
        const float b = sB[k + (lane >> 5)][wn + (lane & 31)];
        accf = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, accf, 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int k = 0; k < PG_K; k += 4) {
        double av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = sA[wm + a * 16 + (lane & 15)][k + (lane >> 4)];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = sB[k + (lane >> 4)][wn + b * 16 + (lane & 15)];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            accd[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], accd[a][b], 0, 0, 0);
      }
    }
  }
  // RAW: theta itself into tT, with the columns up to the next multiple of 64
  // (zero beyond M) written too
  const int Mp = (M + PG_T - 1) / PG_T * PG_T;
  auto store = [&](int64_t sr, int j, T th) {
    if (RAW) {
      if (sr < ns && j < Mp) tT[(int64_t)j * ns + sr] = j < M ? th + (v.b ? v.b[j] : (T)0) : (T)0;
    } else if (sr < ns && j < M) {
      const T tv = tanh_t<T>(th + (v.b ? v.b[j] : (T)0));
      if (t) t[sr * M + j] = tv;
      if (tT) tT[(int64_t)j * ns + sr] = tv;
    }
  };
  if constexpr (sizeof(T) == 4) {
    // element e of lane l: row (e/4)*8 + (l/32)*4 + e%4, col l%32
#pragma unroll
    for (int e = 0; e < 16; ++e)
      store(s0 + wm + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3), j0 + wn + (lane & 31), (T)accf[e]);
  } else {
    // element e of lane l in a 16x16 block: row l/16 + 4e, col l%16
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          store(s0 + wm + a * 16 + (lane >> 4) + 4 * e, j0 + wn + b * 16 + (lane & 15), (T)accd[a][b][e]);
  }
}

template <int JL, class T>
static void launch_prepare(const nka_rbm& m, const int8_t* sigma, int64_t ns, T* t, T* tT,
                           hipStream_t s) {
  const int waves = 4;
  const bool stage = stage_fits<T>(m, waves);
  if (!stage) {
    dim3 grid((unsigned)((ns + PG_T - 1) / PG_T), (unsigned)((m.n_hidden + PG_T - 1) / PG_T));
    hipLaunchKernelGGL((rbm_prepare_gemm_kernel<T, false>), grid, dim3(256), 0, s, m, sigma, ns, t, tT);
    return;
  }
  const size_t lds = (stage ? pad16((size_t)m.n_visible * m.n_hidden * sizeof(T)) : 0) +
                     waves * pad16(m.n_visible);
  const unsigned g = grid_for(ns, waves, 2048);
  if (stage)
    hipLaunchKernelGGL((rbm_prepare_kernel<T, JL, true>), dim3(g), dim3(64 * waves), lds, s, m,
                       sigma, ns, t, tT);
  else
    hipLaunchKernelGGL((rbm_prepare_kernel<T, JL, false>), dim3(g), dim3(64 * waves), lds, s, m,
                       sigma, ns, t, tT);
}
template <int JL>
static void launch_prepare_f32(const nka_rbm& m, const int8_t* sg, int64_t ns, void* t, void* tT,
                               hipStream_t s) {
  launch_prepare<JL, float>(m, sg, ns, (float*)t, (float*)tT, s);
}
template <int JL>
static void launch_prepare_f64(const nka_rbm& m, const int8_t* sg, int64_t ns, void* t, void* tT,
                               hipStream_t s) {
  launch_prepare<JL, double>(m, sg, ns, (double*)t, (double*)tT, s);
}

// ---------------------------------------------------------------------------
// Ising local energy, tiled over samples and hidden units, for FP32 models
// whose W does not fit in LDS.  For a flip of site i,
//   log psi(x') - log psi(x) = -2 x_i a_i + sum_j [lc(t_ij) - lc(theta_j)],
//   t_ij = theta_j - 2 x_i W_ij,  lc(t) = |t| + log(1 + e^{-2|t|}) - log 2,
// so with t' = kappa t (kappa = 2 / ln 2) one unit costs an FMA, the change of
// |t'|, one exp2 and a product update:
//   sum_j [lc(t_j) - lc(theta_j)] = (1/kappa) sum_j (|t'_j| - |theta'_j|)
//                                  + ln2 log2( prod_j (1 + 2^-|t'_j|) / P0 ),
// with P0 the same product at theta (< 2^64 over 64 units).  A workgroup takes
// 64 samples (one per lane) x 256 units (64 per wave, theta' in registers) and
// walks all sites; W' rows are wave-uniform scalar loads.  The four waves'
// partial sums are added in LDS in a fixed order per block of sites and
// written as part[unit tile][site][sample]; a second kernel adds the tiles,
// exponentiates and adds the diagonal.  W is read once per 64 samples instead
// of once per sample.
// ---------------------------------------------------------------------------
constexpr int IF_IB = 8;          // sites per reduction block
constexpr int IF_JT = 256;        // hidden units per workgroup
constexpr float kKappa = 2.8853900817779268f;    // 2 / ln 2
constexpr float kLn2f = 0.69314718055994531f;

__global__ void __launch_bounds__(256) ising_scale_w_kernel(const float* __restrict__ W, int N, int M,
                                                            int Mp, float* __restrict__ Wp) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)N * Mp) return;
  const int i = (int)(e / Mp), j = (int)(e % Mp);
  Wp[e] = j < M ? kKappa * W[(int64_t)i * M + j] : 0.0f;
}

__global__ void __launch_bounds__(256) ising_flip_tile_kernel(
    const float* __restrict__ thT, const float* __restrict__ Wp, const int8_t* __restrict__ sigma,
    int64_t ns, int N, int Mp, const float* __restrict__ a, nka_ising is, int want_lp,
    float* __restrict__ part, double* __restrict__ diag) {
  extern __shared__ uint32_t bits[];             // [words][64]: bit i%32 of word i/32 = (x_i > 0)
  __shared__ float red[2][4][IF_IB][64];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t s0 = (int64_t)blockIdx.x * 64;
  const int jt = blockIdx.y;
  const int jc = jt * IF_JT + w * 64;
  const int nw = (N + 31) >> 5;
  for (int e = tid; e < nw * 64; e += 256) {
    const int r = e & 63, k = e >> 6;
    uint32_t word = 0;
    if (s0 + r < ns) {
      const int8_t* row = sigma + (s0 + r) * N + 32 * k;
      const int lim = min(32, N - 32 * k);
      for (int b = 0; b < lim; ++b) word |= (row[b] > 0 ? 1u : 0u) << b;
    }
    bits[k * 64 + r] = word;
  }
  __syncthreads();

  const int64_t s = min(s0 + lane, ns - 1);
  const bool active = jc < Mp;                   // wave-uniform
  float th[64];
  float rP0 = 1.0f, lcs = 0.0f;
  if (active) {
#pragma unroll
    for (int u = 0; u < 64; ++u) th[u] = thT[(int64_t)(jc + u) * ns + s];
    float P0 = 1.0f;
#pragma unroll
    for (int u = 0; u < 64; ++u) {
      if (want_lp) lcs += log_cosh<float>(th[u]);
      th[u] *= kKappa;
      P0 = fmaf(P0, __builtin_amdgcn_exp2f(-fabsf(th[u])), P0);
    }
    rP0 = 1.0f / P0;
  }

  int buf = 0;
  for (int ib = 0; ib < N; ib += IF_IB) {
    for (int ii = 0; ii < IF_IB; ++ii) {
      const int i = ib + ii;
      float val = 0.0f;
      if (active && i < N) {
        const float g = ((bits[(i >> 5) * 64 + lane] >> (i & 31)) & 1u) ? -2.0f : 2.0f;
        const float* wr = Wp + (int64_t)i * Mp + jc;
        float A = 0.0f, P = 1.0f;
#pragma unroll
        for (int u = 0; u < 64; ++u) {
          const float t = fmaf(g, wr[u], th[u]);
          A += fabsf(t) - fabsf(th[u]);
          P = fmaf(P, __builtin_amdgcn_exp2f(-fabsf(t)), P);
        }
        val = A * (1.0f / kKappa) + kLn2f * __builtin_amdgcn_logf(P * rP0);
      }
      red[buf][w][ii][lane] = val;
    }
    __syncthreads();
    for (int e = tid; e < IF_IB * 64; e += 256) {
      const int ii = e >> 6, r = e & 63, i = ib + ii;
      if (i < N && s0 + r < ns) {
        float v = ((red[buf][0][ii][r] + red[buf][1][ii][r]) + red[buf][2][ii][r]) + red[buf][3][ii][r];
        if (jt == 0 && a) v += ((bits[(i >> 5) * 64 + r] >> (i & 31)) & 1u) ? -2.0f * a[i] : 2.0f * a[i];
        part[((int64_t)jt * (N + 1) + i) * ns + s0 + r] = v;
      }
    }
    buf ^= 1;
  }
  // row N: the tile's share of log psi (sum of log_cosh(theta), plus the
  // visible term in tile 0); tile 0 also writes the diagonal element
  red[buf][w][0][lane] = lcs;
  __syncthreads();
  if (tid < 64 && s0 + tid < ns) {
    const int r = tid;
    if (want_lp) {
      float v = ((red[buf][0][0][r] + red[buf][1][0][r]) + red[buf][2][0][r]) + red[buf][3][0][r];
      if (jt == 0 && a) {
        float vis = 0.0f;
        for (int i = 0; i < N; ++i) vis += ((bits[(i >> 5) * 64 + r] >> (i & 31)) & 1u) ? a[i] : -a[i];
        v += vis;
      }
      part[((int64_t)jt * (N + 1) + N) * ns + s0 + r] = v;
    }
    if (jt == 0) {
      int zz = 0;
      for (int e = 0; e < is.n_edges; ++e) {
        const int i0 = is.edges[2 * e], i1 = is.edges[2 * e + 1];
        const uint32_t b0 = (bits[(i0 >> 5) * 64 + r] >> (i0 & 31)) & 1u;
        const uint32_t b1 = (bits[(i1 >> 5) * 64 + r] >> (i1 & 31)) & 1u;
        zz += b0 == b1 ? 1 : -1;
      }
      diag[s0 + r] = is.mel_dtype == NKA_F64 ? is.J * (double)zz : (double)((float)is.J * (float)zz);
    }
  }
}

__global__ void __launch_bounds__(256) ising_flip_finish_kernel(const float* __restrict__ part,
                                                                const double* __restrict__ diag,
                                                                int64_t ns, int N, int n_jt, double mh,
                                                                double* __restrict__ eloc,
                                                                float* __restrict__ log_psi) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= ns) return;
  double E = diag[s];
  for (int i = 0; i < N; ++i) {
    double d = 0.0;
    for (int jt = 0; jt < n_jt; ++jt) d += (double)part[((int64_t)jt * (N + 1) + i) * ns + s];
    E += mh * exp(d);
  }
  eloc[s] = E;
  if (log_psi) {
    double l = 0.0;
    for (int jt = 0; jt < n_jt; ++jt) l += (double)part[((int64_t)jt * (N + 1) + N) * ns + s];
    log_psi[s] = (float)l;
  }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The tiled path serves FP32 Ising operators with a transverse field on models
// whose W does not fit in LDS.
static bool ising_tiled(const nka_rbm& m, int op_kind, const nka_ising* is) {
  return op_kind == NKA_OP_ISING && is && is->h != 0.0 && m.dtype == NKA_F32 &&
         !stage_fits<float>(m, 4) && (size_t)(m.n_visible + 31) / 32 * 64 * 4 <= 32 * 1024;
}

struct IsingTiles {
  int Mp, n_jt;
  size_t w_bytes, per_sample;
};
static IsingTiles ising_tiles(const nka_rbm& m) {
  IsingTiles t;
  t.Mp = (m.n_hidden + 63) / 64 * 64;
  t.n_jt = (t.Mp + IF_JT - 1) / IF_JT;
  t.w_bytes = align256((size_t)m.n_visible * t.Mp * sizeof(float));
  t.per_sample = ((size_t)t.Mp + (size_t)t.n_jt * (m.n_visible + 1)) * sizeof(float) + sizeof(double);
  return t;
}

static int ising_tiled_eloc(const nka_rbm& m, const nka_ising& is, const int8_t* sigma, int64_t n,
                            double* eloc, float* log_psi, unsigned char* ws, int64_t ws_bytes,
                            hipStream_t s) {
  const IsingTiles tl = ising_tiles(m);
  const int N = m.n_visible, M = m.n_hidden;
  const int64_t avail = ws_bytes - (int64_t)tl.w_bytes - 3 * 256;
  int64_t chunk = avail > 0 ? avail / (int64_t)tl.per_sample : 0;
  if (chunk < n) chunk = chunk / 64 * 64;
  if (chunk > n) chunk = n;
  if (chunk < 1) {
    set_error("nka_rbm_local_energy: workspace too small (%lld bytes, need %lld)", (long long)ws_bytes,
              (long long)(tl.w_bytes + 3 * 256 + 64 * tl.per_sample));
    return NKA_ERR_ARG;
  }
  const RbmView<float> v = rbm_view<float>(m);
  float* Wp = (float*)ws;
  float* thT = (float*)(ws + tl.w_bytes);
  float* part = (float*)(ws + tl.w_bytes + align256((size_t)tl.Mp * chunk * sizeof(float)));
  double* diag = (double*)((unsigned char*)part +
                           align256((size_t)tl.n_jt * (N + 1) * chunk * sizeof(float)));
  const int64_t nwp = (int64_t)N * tl.Mp;
  hipLaunchKernelGGL(ising_scale_w_kernel, dim3((unsigned)((nwp + 255) / 256)), dim3(256), 0, s, v.W,
                     N, M, tl.Mp, Wp);
  const double mh = is.mel_dtype == NKA_F64 ? -is.h : (double)(-(float)is.h);
  const size_t bits_lds = (size_t)(N + 31) / 32 * 64 * sizeof(uint32_t);
  for (int64_t c0 = 0; c0 < n; c0 += chunk) {
    const int64_t nc = min(chunk, n - c0);
    dim3 gg((unsigned)((nc + PG_T - 1) / PG_T), (unsigned)(tl.Mp / PG_T));
    hipLaunchKernelGGL((rbm_prepare_gemm_kernel<float, true>), gg, dim3(256), 0, s, m, sigma + c0 * N,
                       nc, (float*)nullptr, thT);
    dim3 gf((unsigned)((nc + 63) / 64), (unsigned)tl.n_jt);
    hipLaunchKernelGGL(ising_flip_tile_kernel, gf, dim3(256), bits_lds, s, thT, Wp, sigma + c0 * N, nc, N,
                       tl.Mp, v.a, is, log_psi ? 1 : 0, part, diag);
    hipLaunchKernelGGL(ising_flip_finish_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s,
                       part, diag, nc, N, tl.n_jt, mh, eloc + c0, log_psi ? log_psi + c0 : nullptr);
  }
  return NKA_OK;
}

}  // namespace nka

using namespace nka;

extern "C" {

int nka_rbm_log_psi(const nka_rbm* m, const int8_t* sigma, int64_t n, void* out, void* stream) {
  int st = check_rbm(m, "nka_rbm_log_psi");
  if (st) return st;
  NKA_CHECK_ARG(n >= 0 && (n == 0 || (sigma && out)), "nka_rbm_log_psi: bad buffers");
  if (n == 0) return NKA_OK;
  const int jl = pick_jl(m->n_hidden);
  hipStream_t s = (hipStream_t)stream;
  if (m->dtype == NKA_F32) {
    NKA_DISPATCH_JL(jl, launch_log_psi_f32, *m, sigma, n, out, s);
  } else {
    NKA_DISPATCH_JL(jl, launch_log_psi_f64, *m, sigma, n, out, s);
  }
  NKA_CHECK_LAUNCH("nka_rbm_log_psi");
  return NKA_OK;
}

int64_t nka_rbm_local_energy_workspace(const nka_rbm* m, int op_kind, const void* op, int64_t n) {
  if (!m || !op || n <= 0 || check_rbm(m, "nka_rbm_local_energy_workspace") != NKA_OK) return 0;
  if (!ising_tiled(*m, op_kind, op_kind == NKA_OP_ISING ? (const nka_ising*)op : nullptr)) return 0;
  const IsingTiles tl = ising_tiles(*m);
  return (int64_t)(tl.w_bytes + 3 * 256 + (size_t)n * tl.per_sample);
}

int nka_rbm_local_energy(const nka_rbm* m, int op_kind, const void* op, const int8_t* sigma,
                         int64_t n, double* eloc, void* log_psi, void* workspace,
                         int64_t workspace_bytes, void* stream) {
  int st = check_rbm(m, "nka_rbm_local_energy");
  if (st) return st;
  NKA_CHECK_ARG(op != nullptr, "nka_rbm_local_energy: null operator");
  NKA_CHECK_ARG(n >= 0 && (n == 0 || (sigma && eloc)), "nka_rbm_local_energy: bad buffers");
  OpArgs a{};
  a.kind = op_kind;
  if (op_kind == NKA_OP_ISING) {
    a.ising = *(const nka_ising*)op;
    NKA_CHECK_ARG(a.ising.n_edges >= 0 && (a.ising.n_edges == 0 || a.ising.edges),
                  "nka_rbm_local_energy: bad Ising edges");
  } else if (op_kind == NKA_OP_LOCAL) {
    a.local = *(const nka_localop*)op;
    NKA_CHECK_ARG(a.local.n_ops >= 0, "nka_rbm_local_energy: bad local operator");
  } else {
    set_error("nka_rbm_local_energy: unknown operator kind %d", op_kind);
    return NKA_ERR_ARG;
  }
  if (n == 0) return NKA_OK;
  hipStream_t s = (hipStream_t)stream;
  if (ising_tiled(*m, op_kind, &a.ising)) {
    NKA_CHECK_ARG(workspace != nullptr, "nka_rbm_local_energy: this model and operator need a workspace");
    st = ising_tiled_eloc(*m, a.ising, sigma, n, eloc, (float*)log_psi, (unsigned char*)workspace,
                          workspace_bytes, s);
    if (st) return st;
    NKA_CHECK_LAUNCH("nka_rbm_local_energy");
    return NKA_OK;
  }
  const int jl = pick_jl(m->n_hidden);
  if (m->dtype == NKA_F32) {
    NKA_DISPATCH_JL(jl, launch_eloc_f32, *m, a, sigma, n, eloc, log_psi, s, (const nka_eloc_job*)nullptr, 0);
  } else {
    NKA_DISPATCH_JL(jl, launch_eloc_f64, *m, a, sigma, n, eloc, log_psi, s, (const nka_eloc_job*)nullptr, 0);
  }
  NKA_CHECK_LAUNCH("nka_rbm_local_energy");
  return NKA_OK;
}

int nka_rbm_local_energy_batchable(const nka_rbm* shape, const nka_ising* op) {
  if (!shape || !op) return 0;
  return ising_tiled(*shape, NKA_OP_ISING, op) ? 0 : 1;
}

int nka_rbm_local_energy_batched(const nka_eloc_job* jobs, int n_jobs, const nka_rbm* shape,
                                 const nka_ising* op_shape, int64_t n, void* stream) {
  NKA_CHECK_ARG(jobs && shape && op_shape && n_jobs > 0 && n_jobs <= 65535 && n > 0,
                "nka_rbm_local_energy_batched: bad batch");
  NKA_CHECK_ARG(shape->n_visible > 0 && shape->n_hidden > 0 && shape->n_hidden <= kMaxJL * NKA_WAVE &&
                    (shape->dtype == NKA_F32 || shape->dtype == NKA_F64),
                "nka_rbm_local_energy_batched: unsupported model");
  if (ising_tiled(*shape, NKA_OP_ISING, op_shape)) {
    set_error("nka_rbm_local_energy_batched: this model takes the tiled path (W beyond the LDS budget); "
              "call nka_rbm_local_energy per job");
    return NKA_ERR_UNSUPPORTED;
  }
  OpArgs a{};
  a.kind = NKA_OP_ISING;
  a.ising = *op_shape;
  const int jl = pick_jl(shape->n_hidden);
  hipStream_t s = (hipStream_t)stream;
  if (shape->dtype == NKA_F32) {
    NKA_DISPATCH_JL(jl, launch_eloc_f32, *shape, a, (const int8_t*)nullptr, n, (double*)nullptr, (void*)nullptr,
                    s, jobs, n_jobs);
  } else {
    NKA_DISPATCH_JL(jl, launch_eloc_f64, *shape, a, (const int8_t*)nullptr, n, (double*)nullptr, (void*)nullptr,
                    s, jobs, n_jobs);
  }
  NKA_CHECK_LAUNCH("nka_rbm_local_energy_batched");
  return NKA_OK;
}

int nka_rbm_prepare(const nka_rbm* m, const int8_t* sigma, int64_t ns, void* t, void* tT,
                    void* stream) {
  int st = check_rbm(m, "nka_rbm_prepare");
  if (st) return st;
  NKA_CHECK_ARG(ns >= 0 && (ns == 0 || sigma), "nka_rbm_prepare: bad buffers");
  if (ns == 0) return NKA_OK;
  const int jl = pick_jl(m->n_hidden);
  hipStream_t s = (hipStream_t)stream;
  if (m->dtype == NKA_F32) {
    NKA_DISPATCH_JL(jl, launch_prepare_f32, *m, sigma, ns, t, tT, s);
  } else {
    NKA_DISPATCH_JL(jl, launch_prepare_f64, *m, sigma, ns, t, tT, s);
  }
  NKA_CHECK_LAUNCH("nka_rbm_prepare");
  return NKA_OK;
}

int64_t nka_rbm_contract_workspace(const nka_rbm* m, int64_t ns, int n_w) {
  if (!m || ns <= 0 || n_w <= 0) return 0;
  return contract_chunks(*m, ns) * n_w * (n_params(*m) + 1) * (int64_t)sizeof(double);
}

int nka_rbm_contract(const nka_rbm* m, const int8_t* sigma, const void* t, int64_t ns,
                     const double* w, int n_w, const double* mean, double* out, void* workspace,
                     void* stream) {
  int st = check_rbm(m, "nka_rbm_contract");
  if (st) return st;
  NKA_CHECK_ARG(n_w >= 1 && n_w <= kContractMaxW, "nka_rbm_contract: n_w must be 1 or 2");
  NKA_CHECK_ARG(ns > 0 && sigma && t && w && out && workspace, "nka_rbm_contract: bad buffers");
  hipStream_t s = (hipStream_t)stream;
  const int64_t np = n_params(*m);
  const int nsub = contract_nsub(*m, ns);
  const int chunks = (int)contract_chunks(*m, ns);
  dim3 grid(chunks, (m->n_hidden + 63) / 64);
  double* ws = (double*)workspace;
  if (m->dtype == NKA_F32) {
    if (n_w == 1)
      hipLaunchKernelGGL((rbm_contract_partial<float, 1>), grid, dim3(256), 0, s, *m, sigma,
                         (const float*)t, ns, w, ws, np, nsub);
    else
      hipLaunchKernelGGL((rbm_contract_partial<float, 2>), grid, dim3(256), 0, s, *m, sigma,
                         (const float*)t, ns, w, ws, np, nsub);
  } else {
    if (n_w == 1)
      hipLaunchKernelGGL((rbm_contract_partial<double, 1>), grid, dim3(256), 0, s, *m, sigma,
                         (const double*)t, ns, w, ws, np, nsub);
    else
      hipLaunchKernelGGL((rbm_contract_partial<double, 2>), grid, dim3(256), 0, s, *m, sigma,
                         (const double*)t, ns, w, ws, np, nsub);
  }
  const int64_t len = (int64_t)n_w * np;
  unsigned g = (unsigned)std::min<int64_t>((len + 255) / 256, 4096);
  hipLaunchKernelGGL(contract_reduce, dim3(g), dim3(256), 0, s, (const double*)workspace, chunks,
                     n_w, np, mean, out);
  NKA_CHECK_LAUNCH("nka_rbm_contract");
  return NKA_OK;
}

int nka_rbm_jacobian(const nka_rbm* m, const int8_t* sigma, const void* t, const void* tT,
                     int64_t ns, const double* mean, double scale, int layout, int out_kind,
                     void* out_a, void* out_b, int64_t ldo, void* stream) {
  int st = check_rbm(m, "nka_rbm_jacobian");
  if (st) return st;
  NKA_CHECK_ARG(ns > 0 && sigma && mean && out_a, "nka_rbm_jacobian: bad buffers");
  NKA_CHECK_ARG(out_kind == NKA_OUT_F32 || out_kind == NKA_OUT_F64 ||
                    (out_kind == NKA_OUT_BF16X2 && out_b),
                "nka_rbm_jacobian: bad output kind");
  NKA_CHECK_ARG(layout == NKA_LAYOUT_PS ? tT != nullptr : (layout == NKA_LAYOUT_SP && t),
                "nka_rbm_jacobian: bad layout / missing t");
  hipStream_t s = (hipStream_t)stream;
  const int64_t np = n_params(*m);
  NKA_CHECK_ARG(ldo >= (layout == NKA_LAYOUT_PS ? ns : np), "nka_rbm_jacobian: ldo too small");
  const int64_t total = np * ns;
  const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
  if (layout == NKA_LAYOUT_PS) {
    // 16-byte runs of the bf16 planes need 8-aligned rows of the output
    NKA_CHECK_ARG(out_kind != NKA_OUT_BF16X2 || (ldo % 8 == 0 && (uintptr_t)out_a % 16 == 0 &&
                                                 (uintptr_t)out_b % 16 == 0),
                  "nka_rbm_jacobian: bf16 planes need ldo %% 8 == 0 and 16-byte alignment");
    const dim3 grid((unsigned)((ns + kJacChunk - 1) / kJacChunk), (unsigned)(m->n_visible + 2));
    const bool vec = ns % 8 == 0 && (uintptr_t)tT % 16 == 0;
    if (m->dtype == NKA_F32) {
      if (vec)
        hipLaunchKernelGGL((rbm_jac_ps_kernel<float, true>), grid, dim3(256), 0, s, *m, sigma,
                           (const float*)tT, ns, mean, scale, out_kind, out_a, out_b, np, ldo);
      else
        hipLaunchKernelGGL((rbm_jac_ps_kernel<float, false>), grid, dim3(256), 0, s, *m, sigma,
                           (const float*)tT, ns, mean, scale, out_kind, out_a, out_b, np, ldo);
    } else {
      if (vec)
        hipLaunchKernelGGL((rbm_jac_ps_kernel<double, true>), grid, dim3(256), 0, s, *m, sigma,
                           (const double*)tT, ns, mean, scale, out_kind, out_a, out_b, np, ldo);
      else
        hipLaunchKernelGGL((rbm_jac_ps_kernel<double, false>), grid, dim3(256), 0, s, *m, sigma,
                           (const double*)tT, ns, mean, scale, out_kind, out_a, out_b, np, ldo);
    }
  } else {
    if (m->dtype == NKA_F32)
      hipLaunchKernelGGL(rbm_jac_sp_kernel<float>, dim3(g), dim3(256), 0, s, *m, sigma,
                         (const float*)t, ns, mean, scale, out_kind, out_a, out_b, np, ldo);
    else
      hipLaunchKernelGGL(rbm_jac_sp_kernel<double>, dim3(g), dim3(256), 0, s, *m, sigma,
                         (const double*)t, ns, mean, scale, out_kind, out_a, out_b, np, ldo);
  }
  NKA_CHECK_LAUNCH("nka_rbm_jacobian");
  return NKA_OK;
}

}  // extern "C"
