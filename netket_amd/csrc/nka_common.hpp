// Shared device helpers for the gfx950 VMC-step kernels.
//
// * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
//   "Parallel random numbers: as easy as 1, 2, 3", SC'11), used for every
//   random draw of the sampler so that a draw is a pure function of
//   (seed, chain, proposal index).
// * log_cosh(x) = |x| + log1p(exp(-2|x|)) - log 2, the stable form used by
//   the reference activation (netket/nn/activation.py:78-84).
// * 64-lane wavefront reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#define NKA_WAVE 64

namespace nka {

struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c.z;
    u32x4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
  }
  return c;
}

// Stream tags placed in the fourth counter word so that different uses of the
// generator never share a counter.
enum : uint32_t {
  TAG_METROPOLIS = 0x4D455452u,  // "METR"
  TAG_RANDSTATE = 0x52414E44u,   // "RAND"
};

// Uniform in [0, 1) from the top 24 bits (float) or 53 bits (double).
template <class T> __device__ inline T uniform01(const u32x4& r);
template <> __device__ inline float uniform01<float>(const u32x4& r) {
  return (float)(r.y >> 8) * 5.9604644775390625e-08f;  // 2^-24
}
template <> __device__ inline double uniform01<double>(const u32x4& r) {
  const uint64_t u = (((uint64_t)r.y << 32) | (uint64_t)r.z) >> 11;
  return (double)u * 1.1102230246251565e-16;  // 2^-53
}

// floor(x * n / 2^32): an index in [0, n) from one 32-bit draw.
__host__ __device__ inline uint32_t bounded(uint32_t x, uint32_t n) {
  return (uint32_t)(((uint64_t)x * (uint64_t)n) >> 32);
}

template <class T> __device__ inline T log_cosh(T x);
// FP32: hardware exp2/log2 (v_exp_f32, v_log_f32).  With e = exp(-2|x|) in
// (0, 1], log(1 + e) is evaluated directly; the rounding of 1 + e costs at
// most 2^-24 absolute, below the FP32 resolution of the result.
template <> __device__ inline float log_cosh<float>(float x) {
  const float ax = fabsf(x);
  const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * ax);  // exp(-2|x|)
  return ax + 0.69314718055994531f * (__builtin_amdgcn_logf(1.0f + e) - 1.0f);
}
template <> __device__ inline double log_cosh<double>(double x) {
  const double ax = fabs(x);
  return ax + log1p(exp(-2.0 * ax)) - 0.69314718055994530942;
}

template <class T> __device__ inline T exp_t(T x);
template <> __device__ inline float exp_t<float>(float x) { return __expf(x); }
template <> __device__ inline double exp_t<double>(double x) { return exp(x); }

template <class T> __device__ inline T log_t(T x);
template <> __device__ inline float log_t<float>(float x) { return logf(x); }
template <> __device__ inline double log_t<double>(double x) { return log(x); }

template <class T> __device__ inline T tanh_t(T x);
template <> __device__ inline float tanh_t<float>(float x) { return tanhf(x); }
template <> __device__ inline double tanh_t<double>(double x) { return tanh(x); }

// Sum over the 64 lanes of a wavefront; every lane gets the total.
// Generic: butterfly over cross-lane shuffles.
template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, NKA_WAVE);
  return v;
}

// FP32: DPP row operations inside 16-lane rows, then row broadcasts, so the
// partial sums never leave the VALU; the total lands in lane 63 and is read
// back as a scalar.  Lanes masked off by a DPP step receive 0.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ inline float dpp_f32(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, BANK_MASK, false));
}
template <> __device__ inline float wave_sum<float>(float v) {
  v += dpp_f32<0xb1, 0xf, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_f32<0x4e, 0xf, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_f32<0x114, 0xf, 0xe>(v);  // row_shr:4
  v += dpp_f32<0x118, 0xf, 0xc>(v);  // row_shr:8
  v += dpp_f32<0x142, 0xa, 0xf>(v);  // row_bcast:15
  v += dpp_f32<0x143, 0xc, 0xf>(v);  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline int lane_id() { return (int)(threadIdx.x & (NKA_WAVE - 1)); }

// Position of the k-th (0-based) set bit of m (m has more than k set bits).
__device__ inline int kth_set_bit(uint64_t m, int k) {
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) {
    const int c = __popcll(m & ((1ull << w) - 1ull));
    if (k >= c) {
      k -= c;
      m >>= w;
      pos += w;
    }
  }
  return pos;
}

// Sum of a 32-bit integer over the wavefront, in every lane (exact).
__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
  return v;
}

// Weighted choice among the clusters a wavefront holds in slots (cluster
// lane + 64 r carries integer weight w[r], 0 when it cannot be picked): the
// cluster whose cumulative weight, in cluster order, first exceeds k
// (0 <= k < the total).  Integer sums, so the choice is exact; -1 if none.
template <int R>
__device__ inline int weighted_pick(const uint32_t (&w)[R], int RS, uint32_t k) {
  const int lane = lane_id();
  int sel = -1;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (sel < 0 && r < RS) {
      const uint32_t tot = wave_sum_u32(w[r]);
      if (k < tot) {
        uint32_t inc = w[r];   // inclusive prefix sum over the lanes
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t t = (uint32_t)__shfl_up((int)inc, off);
          if (lane >= off) inc += t;
        }
        const uint64_t over = __ballot(inc > k);
        sel = r * 64 + (int)__builtin_ctzll(over);
      } else {
        k -= tot;
      }
    }
  }
  return sel;
}

// The two sites of cluster sel (wave-uniform) from the lanes' register slots
// (cluster lane + 64 r in slot r): a lane read instead of a dependent global
// load on the proposal's critical path.
template <int R>
__device__ inline void cluster_sites(const int (&cu)[R], const int (&cv)[R], int sel, int& s0, int& s1) {
  const int rs = sel >> 6, ln = sel & 63;
  s0 = 0;
  s1 = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r == rs) {
      s0 = __builtin_amdgcn_readlane(cu[r], ln);
      s1 = __builtin_amdgcn_readlane(cv[r], ln);
    }
  }
}

// Order this lane's LDS writes before later LDS reads of other lanes in the
// same wavefront.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane 0's value in every lane (a scalar register), so that decisions taken on
// it are wave-uniform by construction.
template <class T> __device__ inline T uniform_value(T v);
template <> __device__ inline int uniform_value<int>(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}
template <> __device__ inline uint32_t uniform_value<uint32_t>(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
template <> __device__ inline float uniform_value<float>(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
template <> __device__ inline double uniform_value<double>(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// FP32 -> bf16, round to nearest even (NaN stays NaN); and back.
__device__ inline uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Two FP32 values rounded to nearest even bf16 by the hardware conversion
// (v_cvt_pk_bf16_f32), packed low-first; for finite values the same bits as
// bf16_rne.  And the hi + lo split of two values: hi = bf16(f),
// lo = bf16(f - hi), each pair packed low-first.
typedef float nka_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 nka_bf16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t bf16x2_rne(float a, float b) {
  const nka_f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, nka_bf16x2));
}
__device__ inline void bf16x2_split(float a, float b, uint32_t& hi, uint32_t& lo) {
  hi = bf16x2_rne(a, b);
  lo = bf16x2_rne(a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xFFFF0000u));
}

// Lower-triangular tile pair (ti >= tj) of linear id b in an order made of
// SxS super-blocks of tiles (row-major over the super-blocks of the lower
// triangle, then row-major inside each; diagonal super-blocks are
// triangular).  T is the number of tile rows.  Tiles that run at the same time
// on one XCD then share a few operand panels (about S rows + S columns)
// instead of one row panel and a different column panel each.
__device__ inline void tri_supertile(int64_t b, int64_t T, int& ti, int& tj) {
  constexpr int64_t S = 8;
  const int64_t SB = (T + S - 1) / S;
  int64_t rem = b;
  for (int64_t I = 0; I < SB; ++I) {
    const int64_t rows = min<int64_t>(S, T - I * S);
    for (int64_t J = 0; J <= I; ++J) {
      const int64_t cols = min<int64_t>(S, T - J * S);
      const int64_t cnt = (I == J) ? rows * (rows + 1) / 2 : rows * cols;
      if (rem < cnt) {
        int64_t r, c;
        if (I == J) {
          r = (int64_t)((sqrt(8.0 * (double)rem + 1.0) - 1.0) * 0.5);
          while ((r + 1) * (r + 2) / 2 <= rem) ++r;
          while (r * (r + 1) / 2 > rem) --r;
          c = rem - r * (r + 1) / 2;
        } else {
          r = rem / cols;
          c = rem % cols;
        }
        ti = (int)(I * S + r);
        tj = (int)(J * S + c);
        return;
      }
      rem -= cnt;
    }
  }
  ti = tj = 0;  // not reached for b < T (T + 1) / 2
}

}  // namespace nka

// ---------------------------------------------------------------------------
// Host-side status plumbing shared by every C-ABI entry point.
// ---------------------------------------------------------------------------
namespace nka {
void set_error(const char* fmt, ...);
}

#define NKA_OK 0
#define NKA_ERR_ARG -1
#define NKA_ERR_HIP -2
#define NKA_ERR_UNSUPPORTED -3

#define NKA_CHECK_ARG(cond, ...)      \
  do {                                \
    if (!(cond)) {                    \
      nka::set_error(__VA_ARGS__);    \
      return NKA_ERR_ARG;             \
    }                                 \
  } while (0)

#define NKA_CHECK_HIP(expr, what)                                      \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess) {                                            \
      nka::set_error("%s: %s", what, hipGetErrorString(e_));           \
      return NKA_ERR_HIP;                                              \
    }                                                                  \
  } while (0)

#define NKA_CHECK_LAUNCH(what)                                               \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess) {                                                  \
      nka::set_error("%s: launch failed: %s", what, hipGetErrorString(e_)); \
      return NKA_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#define NKA_F32 0
#define NKA_F64 1
