// Wavefront-per-configuration helpers for the RBM ansatz
// log psi(s) = sum_j log_cosh(b_j + sum_i s_i W_ij) + sum_i a_i s_i
// (netket/models/rbm.py:57-81).
//
// Lane l of a wavefront owns hidden units j = l + 64*jj (jj < JL), so the
// hidden pre-activations theta_j live in registers and a change of k visible
// spins costs k W-row reads (contiguous in j) plus JL log_cosh per lane and one
// 64-lane reduction.  W is read from LDS when it fits (staged once per
// workgroup) and from global memory otherwise.
#pragma once

#include "nka_common.hpp"
#include "netket_amd.h"

namespace nka {

// Hidden-unit registers per lane, chosen from the number of hidden units.
inline int pick_jl(int M) {
  const int need = (M + NKA_WAVE - 1) / NKA_WAVE;
  int jl = 1;
  while (jl < need) jl *= 2;
  return jl;
}

constexpr int kMaxJL = 64;            // M <= 4096
constexpr size_t kLdsBudget = 64 * 1024;

template <class T, int JL>
struct HiddenState {
  T th[JL];  // theta_j
  T lc[JL];  // log_cosh(theta_j)
};

// theta_j = b_j + sum_i s_i W_ij for this lane's hidden units.
template <class T, int JL>
__device__ inline void compute_theta(HiddenState<T, JL>& h, const int8_t* xs, const T* W,
                                     const T* bias, int N, int M) {
  const int lane = lane_id();
#pragma unroll
  for (int jj = 0; jj < JL; ++jj) {
    const int j = lane + NKA_WAVE * jj;
    T acc = (T)0;
    if (j < M) {
      acc = bias ? bias[j] : (T)0;
      for (int i = 0; i < N; ++i) {
        const T w = W[(size_t)i * M + j];
        acc += (xs[i] > 0) ? w : -w;
      }
      h.th[jj] = acc;
      h.lc[jj] = log_cosh<T>(acc);
    } else {
      h.th[jj] = (T)0;
      h.lc[jj] = (T)0;
    }
  }
}

// sum_j log_cosh(theta_j) over the wavefront.
template <class T, int JL>
__device__ inline T sum_log_cosh(const HiddenState<T, JL>& h) {
  T s = (T)0;
#pragma unroll
  for (int jj = 0; jj < JL; ++jj) s += h.lc[jj];
  return wave_sum(s);
}

// sum_i a_i s_i over the wavefront (0 without visible bias).
template <class T>
__device__ inline T visible_term(const int8_t* xs, const T* a, int N) {
  if (!a) return (T)0;
  T s = (T)0;
  for (int i = lane_id(); i < N; i += NKA_WAVE) s += (xs[i] > 0) ? a[i] : -a[i];
  return wave_sum(s);
}

// Change of log psi when the spins at sites[0..nf) flip sign.  Fills the new
// theta and log_cosh values into nth / nlc so a caller that accepts the move
// can commit them without recomputation.
template <class T, int JL>
__device__ inline T flip_delta(const HiddenState<T, JL>& h, const int8_t* xs, const T* W,
                               const T* a, int N, int M, const int* sites, int nf, T* nth,
                               T* nlc) {
  const int lane = lane_id();
  T part = (T)0;
#pragma unroll
  for (int jj = 0; jj < JL; ++jj) {
    const int j = lane + NKA_WAVE * jj;
    if (j < M) {
      T d = (T)0;
      for (int f = 0; f < nf; ++f) {
        const int s = sites[f];
        const T w = W[(size_t)s * M + j];
        d += (xs[s] > 0) ? -2 * w : 2 * w;
      }
      const T t = h.th[jj] + d;
      const T l = log_cosh<T>(t);
      nth[jj] = t;
      nlc[jj] = l;
      part += l - h.lc[jj];
    } else {
      nth[jj] = (T)0;
      nlc[jj] = (T)0;
    }
  }
  T delta = wave_sum(part);
  if (a) {
    for (int f = 0; f < nf; ++f) {
      const int s = sites[f];
      delta += (xs[s] > 0) ? -2 * a[s] : 2 * a[s];
    }
  }
  return delta;
}

template <class T>
struct RbmView {
  const T* W;
  const T* b;
  const T* a;
  int N, M;
};

template <class T>
__host__ __device__ inline RbmView<T> rbm_view(const nka_rbm& m) {
  const T* P = (const T*)m.params;
  RbmView<T> v;
  v.W = P + m.kernel_off;
  v.b = m.bias_off >= 0 ? P + m.bias_off : nullptr;
  v.a = m.visible_off >= 0 ? P + m.visible_off : nullptr;
  v.N = m.n_visible;
  v.M = m.n_hidden;
  return v;
}

// Stage W (N x M) into LDS with the whole workgroup.
template <class T>
__device__ inline void stage_w(T* Wl, const T* Wg, int N, int M) {
  const int total = N * M;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) Wl[idx] = Wg[idx];
  __syncthreads();
}

__host__ __device__ inline size_t pad16(size_t x) { return (x + 15) & ~(size_t)15; }

// Dispatch a templated launcher on the hidden-register count.
#define NKA_DISPATCH_JL(jl, FN, ...)            \
  switch (jl) {                                 \
    case 1: FN<1>(__VA_ARGS__); break;          \
    case 2: FN<2>(__VA_ARGS__); break;          \
    case 4: FN<4>(__VA_ARGS__); break;          \
    case 8: FN<8>(__VA_ARGS__); break;          \
    case 16: FN<16>(__VA_ARGS__); break;        \
    case 32: FN<32>(__VA_ARGS__); break;        \
    case 64: FN<64>(__VA_ARGS__); break;        \
    default: break;                             \
  }

}  // namespace nka
