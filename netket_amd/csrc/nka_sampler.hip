// Metropolis-Hastings sampler for the RBM, one wavefront per Markov chain.
//
// Semantics (netket/sampler/metropolis.py:382-505, netket/vqs/mc/mc_state/state.py:521-576):
//   reset: log_prob = machine_pow * Re log psi(sigma)          (metropolis.py:398-403)
//   repeat n_discard + chain_length times a sweep of sweep_size proposals;
//   each proposal sigma' is accepted when u < exp(log_prob' - log_prob + corr)
//   (metropolis.py:441-450); after each of the last chain_length sweeps the
//   chain state is recorded (metropolis.py:462-464, 492-499).
// Rules:
//   LocalRule    : uniform site, spin flipped (rules/local.py:40-49,
//                  hilbert/random/homogeneous.py:146-170)
//   ExchangeRule : a cluster (i, j) uniformly among those with s_i != s_j,
//                  swapped; corr = log n_diff(sigma) - log n_diff(sigma')
//                  (rules/exchange.py:143-218)
// The wavefront keeps theta_j and log_cosh(theta_j) of its chain in registers
// (64 hidden units per register slot) and the spins in LDS, so a proposal
// reads only the W rows of the flipped sites.
#include "nka_rbm.hpp"

namespace nka {

// ---------------------------------------------------------------------------
// Initial states
// ---------------------------------------------------------------------------
__global__ void random_states_kernel(int8_t* __restrict__ sigma, int C, int N, int n_down,
                                     uint32_t k0, uint32_t k1, uint64_t counter) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  int8_t* x = sigma + (int64_t)c * N;
  if (n_down < 0) {
    for (int i0 = 0; i0 < N; i0 += 128) {
      const uint64_t q = counter + (uint64_t)(i0 / 128);
      u32x4 r = philox4x32_10({(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)c, TAG_RANDSTATE}, k0, k1);
      const uint32_t words[4] = {r.x, r.y, r.z, r.w};
      for (int i = i0; i < min(N, i0 + 128); ++i) {
        const int bit = (words[(i - i0) >> 5] >> ((i - i0) & 31)) & 1;
        x[i] = bit ? (int8_t)-1 : (int8_t)1;
      }
    }
  } else {
    // n_down spins -1 at uniformly random positions: Fisher-Yates shuffle of
    // [-1]*n_down + [+1]*(N-n_down), swap partner j uniform in [0, i].
    for (int i = 0; i < N; ++i) x[i] = (i < n_down) ? (int8_t)-1 : (int8_t)1;
    for (int i = N - 1; i > 0; --i) {
      const uint64_t q = counter + (uint64_t)i;
      u32x4 r = philox4x32_10({(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)c, TAG_RANDSTATE}, k0, k1);
      const int j = (int)bounded(r.x, (uint32_t)(i + 1));
      const int8_t tmp = x[i];
      x[i] = x[j];
      x[j] = tmp;
    }
  }
}

// ---------------------------------------------------------------------------
// Sweeps
// ---------------------------------------------------------------------------
constexpr int kMaxClusterSlots = 16;  // clusters per lane: n_clusters <= 1024

struct SweepArgs {
  int rule;
  const int32_t* clusters;
  int n_clusters;
  const uint32_t* weights;      // exchange: integer weight per cluster (NULL: uniform)
  int C, sweep_size, n_discard, chain_length;
  double machine_pow;
  uint32_t k0, k1;
  uint64_t counter;
  const uint64_t* counter_dev;  // when set, the counter is read from device memory
};

template <class T, int JL, bool STAGE>
__device__ __forceinline__ void rbm_metropolis_kernel_body(const nka_rbm& m, const SweepArgs& sa, int8_t* __restrict__ sigma, T* __restrict__ log_prob, int8_t* __restrict__ samples, T* __restrict__ lp_samples, int64_t* __restrict__ n_accepted, int bx) {
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
  const int c = bx * waves + wid;
  if (c >= sa.C) return;

  for (int i = lane; i < N; i += NKA_WAVE) xs[i] = sigma[(int64_t)c * N + i];
  wave_lds_sync();
  HiddenState<T, JL> h;
  compute_theta<T, JL>(h, xs, W, v.b, N, M);
  const T pw = (T)sa.machine_pow;
  T lp = pw * (sum_log_cosh<T, JL>(h) + visible_term<T>(xs, v.a, N));

  // exchange clusters owned by this lane: cl = lane + 64 * r
  const bool exch = sa.rule == NKA_RULE_EXCHANGE;
  const int RS = exch ? (sa.n_clusters + NKA_WAVE - 1) / NKA_WAVE : 0;
  int cu[kMaxClusterSlots], cv[kMaxClusterSlots];
#pragma unroll
  for (int r = 0; r < kMaxClusterSlots; ++r) {
    const int cl = lane + NKA_WAVE * r;
    if (r < RS && cl < sa.n_clusters) {
      cu[r] = sa.clusters[2 * cl];
      cv[r] = sa.clusters[2 * cl + 1];
    } else {
      cu[r] = -1;
      cv[r] = -1;
    }
  }

  int64_t nacc = 0;
  uint64_t q = sa.counter;
  if (sa.counter_dev) {
    const uint64_t cd = *sa.counter_dev;
    q = ((uint64_t)uniform_value<uint32_t>((uint32_t)(cd >> 32)) << 32) |
        (uint64_t)uniform_value<uint32_t>((uint32_t)cd);
  }
  T nth[JL], nlc[JL];
  const int n_sweeps = sa.n_discard + sa.chain_length;
  // the draw of the next proposal is computed one proposal ahead, so its
  // latency overlaps the current proposal's reductions
  u32x4 rnd_next = philox4x32_10({(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)c, TAG_METROPOLIS},
                                 sa.k0, sa.k1);
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int p = 0; p < sa.sweep_size; ++p, ++q) {
      const u32x4 rnd = rnd_next;
      {
        const uint64_t qn = q + 1;
        rnd_next = philox4x32_10({(uint32_t)qn, (uint32_t)(qn >> 32), (uint32_t)c, TAG_METROPOLIS},
                                 sa.k0, sa.k1);
      }
      int sites[2];
      int nf;
      T corr = (T)0;
      bool valid = true;
      if (!exch) {
        sites[0] = (int)bounded(uniform_value<uint32_t>(rnd.x), (uint32_t)N);
        nf = 1;
      } else if (sa.weights) {
        // weighted choice: cluster cl of the differing ones with probability
        // weights[cl] / (their total); correction log total - log total'
        uint32_t wd[kMaxClusterSlots];
        uint32_t tot = 0;
#pragma unroll
        for (int r = 0; r < kMaxClusterSlots; ++r) {
          const bool d = r < RS && cu[r] >= 0 && xs[cu[r]] != xs[cv[r]];
          wd[r] = d ? sa.weights[lane + NKA_WAVE * r] : 0u;
          if (r < RS) tot += wave_sum_u32(wd[r]);
        }
        tot = uniform_value<uint32_t>(tot);
        if (tot == 0) {
          valid = false;
          nf = 0;
        } else {
          const uint32_t k = bounded(uniform_value<uint32_t>(rnd.x), tot);
          int sel = uniform_value<int>(weighted_pick<kMaxClusterSlots>(wd, RS, k));
          if (sel < 0 || sel >= sa.n_clusters) sel = 0;  // unreachable: k < tot
          int su, sv;
          cluster_sites<kMaxClusterSlots>(cu, cv, sel, su, sv);
          sites[0] = su;
          sites[1] = sv;
          nf = 2;
          uint32_t tot_new = 0;
#pragma unroll
          for (int r = 0; r < kMaxClusterSlots; ++r) {
            if (r < RS) {
              bool d = false;
              if (cu[r] >= 0) {
                const int a = cu[r], bb = cv[r];
                const int8_t xa = a == su ? xs[sv] : (a == sv ? xs[su] : xs[a]);
                const int8_t xb = bb == su ? xs[sv] : (bb == sv ? xs[su] : xs[bb]);
                d = xa != xb;
              }
              tot_new += wave_sum_u32(d ? sa.weights[lane + NKA_WAVE * r] : 0u);
            }
          }
          tot_new = uniform_value<uint32_t>(tot_new);
          corr = uniform_value<T>(log_t<T>((T)tot) - log_t<T>((T)tot_new));
        }
      } else {
        // differing clusters of the current state
        uint64_t bal[kMaxClusterSlots];
        int n_diff = 0;
#pragma unroll
        for (int r = 0; r < kMaxClusterSlots; ++r) {
          if (r < RS) {
            const bool d = cu[r] >= 0 && xs[cu[r]] != xs[cv[r]];
            bal[r] = __ballot(d);
            n_diff += __popcll(bal[r]);
          } else {
            bal[r] = 0;
          }
        }
        n_diff = uniform_value<int>(n_diff);
        if (n_diff == 0) {
          valid = false;
          nf = 0;
        } else {
          int k = (int)bounded(uniform_value<uint32_t>(rnd.x), (uint32_t)n_diff);
          int sel = -1;
#pragma unroll
          for (int r = 0; r < kMaxClusterSlots; ++r) {
            if (sel < 0 && r < RS) {
              const int cnt = __popcll(bal[r]);
              if (k < cnt) {
                sel = r * NKA_WAVE + kth_set_bit(bal[r], k);
              } else {
                k -= cnt;
              }
            }
          }
          sel = uniform_value<int>(sel);
          if (sel < 0 || sel >= sa.n_clusters) sel = 0;  // unreachable: k < n_diff
          int su, sv;
          cluster_sites<kMaxClusterSlots>(cu, cv, sel, su, sv);
          sites[0] = su;
          sites[1] = sv;
          nf = 2;
          // differing clusters after the swap (sigma'_u = sigma_v, sigma'_v = sigma_u)
          int n_new = 0;
#pragma unroll
          for (int r = 0; r < kMaxClusterSlots; ++r) {
            if (r < RS) {
              bool d = false;
              if (cu[r] >= 0) {
                const int a = cu[r], bb = cv[r];
                const int8_t xa = a == su ? xs[sv] : (a == sv ? xs[su] : xs[a]);
                const int8_t xb = bb == su ? xs[sv] : (bb == sv ? xs[su] : xs[bb]);
                d = xa != xb;
              }
              n_new += __popcll(__ballot(d));
            }
          }
          n_new = uniform_value<int>(n_new);
          corr = uniform_value<T>(log_t<T>((T)n_diff) - log_t<T>((T)n_new));
        }
      }
      if (valid) {
        const T delta =
            uniform_value<T>(flip_delta<T, JL>(h, xs, W, v.a, N, M, sites, nf, nth, nlc));
        const T u = uniform_value<T>(uniform01<T>(rnd));
        const bool acc = u < exp_t<T>(pw * delta + corr);
        if (acc) {
#pragma unroll
          for (int jj = 0; jj < JL; ++jj) {
            h.th[jj] = nth[jj];
            h.lc[jj] = nlc[jj];
          }
          if (lane == 0) {
            for (int f = 0; f < nf; ++f) xs[sites[f]] = (int8_t)(-xs[sites[f]]);
          }
          lp += pw * delta;
          ++nacc;
          wave_lds_sync();
        }
      }
    }
    if (sw >= sa.n_discard) {
      const int t = sw - sa.n_discard;
      int8_t* dst = samples + ((int64_t)c * sa.chain_length + t) * N;
      for (int i = lane; i < N; i += NKA_WAVE) dst[i] = xs[i];
      if (lp_samples && lane == 0) lp_samples[(int64_t)c * sa.chain_length + t] = lp;
    }
  }
  for (int i = lane; i < N; i += NKA_WAVE) sigma[(int64_t)c * N + i] = xs[i];
  if (lane == 0) {
    log_prob[c] = lp;
    n_accepted[c] += nacc;
  }
}

template <class T, int JL, bool STAGE>
__global__ void __launch_bounds__(256) rbm_metropolis_kernel(nka_rbm m, SweepArgs sa, int8_t* __restrict__ sigma, T* __restrict__ log_prob, int8_t* __restrict__ samples, T* __restrict__ lp_samples, int64_t* __restrict__ n_accepted) {
  rbm_metropolis_kernel_body<T, JL, STAGE>(m, sa, sigma, log_prob, samples, lp_samples, n_accepted, blockIdx.x);
}

// The same sweeps for a batch of jobs of one shape: blockIdx.y selects the job
// (its parameters, chains, outputs and random stream); the shape arguments are shared.
template <class T, int JL, bool STAGE>
__global__ void __launch_bounds__(256) rbm_metropolis_kernel_batched(const nka_sweep_job* __restrict__ jobs, SweepArgs sa) {
  const nka_sweep_job& j = jobs[blockIdx.y];
  SweepArgs s = sa;
  s.k0 = (uint32_t)j.seed;
  s.k1 = (uint32_t)(j.seed >> 32);
  s.counter = j.counter;
  s.counter_dev = j.counter_dev;
  rbm_metropolis_kernel_body<T, JL, STAGE>(j.model, s, j.sigma, (T*)j.log_prob, j.samples, (T*)j.log_prob_samples, j.n_accepted, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Local-rule sweeps of FP32 models whose W does not fit in LDS.  One
// wavefront per chain, lane l owning the units j = l + 64 jj, as in
// rbm_metropolis_kernel, with two changes for a W that streams from the
// caches.  The site of a local-rule proposal does not depend on the state, so
// the W row of the next proposal is loaded one proposal ahead and its latency
// overlaps the current proposal.  And the change of log psi uses
//   sum_j [lc(t_j) - lc(theta_j)] = (1/kappa) sum_j (|t'_j| - |theta'_j|)
//                                  + ln2 log2( prod_j (1 + 2^-|t'_j|) / P0 )
// (lc(t) = |t| + log(1 + e^{-2|t|}) - log 2, t' = kappa t, kappa = 2/ln 2):
// one exp2 per unit instead of an exp and a log.  Each lane keeps theta' of
// its units and 1/P0 of its product (< 2^64 over at most 64 units); an
// accepted move replaces both by the proposal's.  Same random streams and
// acceptance rule as rbm_metropolis_kernel.
// ---------------------------------------------------------------------------
constexpr float kKappaS = 2.8853900817779268f;    // 2 / ln 2
constexpr float kLn2S = 0.69314718055994531f;

template <int JL>
__device__ __forceinline__ void rbm_metropolis_wide_kernel_body(const nka_rbm& m, const SweepArgs& sa, int8_t* __restrict__ sigma, float* __restrict__ log_prob, int8_t* __restrict__ samples, float* __restrict__ lp_samples, int64_t* __restrict__ n_accepted, int bx) {
  extern __shared__ unsigned char smem[];
  const RbmView<float> v = rbm_view<float>(m);
  const int N = v.N, M = v.M;
  const int wid = threadIdx.x / NKA_WAVE, waves = blockDim.x / NKA_WAVE;
  const int lane = lane_id();
  int8_t* xs = (int8_t*)smem + wid * pad16(N);
  const int c = bx * waves + wid;
  if (c >= sa.C) return;
  for (int i = lane; i < N; i += NKA_WAVE) xs[i] = sigma[(int64_t)c * N + i];
  wave_lds_sync();

  // theta of this lane's units: all JL loads of a row in flight per site
  float th[JL];
#pragma unroll
  for (int jj = 0; jj < JL; ++jj) {
    const int j = lane + NKA_WAVE * jj;
    th[jj] = (j < M && v.b) ? v.b[j] : 0.0f;
  }
  for (int i = 0; i < N; ++i) {
    const bool up = xs[i] > 0;
    const float* Wr = v.W + (size_t)i * M;
#pragma unroll
    for (int jj = 0; jj < JL; ++jj) {
      const int j = lane + NKA_WAVE * jj;
      const float w = j < M ? Wr[j] : 0.0f;
      th[jj] += up ? w : -w;
    }
  }
  const float pw = (float)sa.machine_pow;
  float lcs = 0.0f, P0 = 1.0f;
#pragma unroll
  for (int jj = 0; jj < JL; ++jj) {
    lcs += lane + NKA_WAVE * jj < M ? log_cosh<float>(th[jj]) : 0.0f;
    th[jj] *= kKappaS;
    P0 = fmaf(P0, __builtin_amdgcn_exp2f(-fabsf(th[jj])), P0);
  }
  float rP0 = 1.0f / P0;
  float lp = pw * (wave_sum(lcs) + visible_term<float>(xs, v.a, N));

  int64_t nacc = 0;
  uint64_t q = sa.counter;
  if (sa.counter_dev) {
    const uint64_t cd = *sa.counter_dev;
    q = ((uint64_t)uniform_value<uint32_t>((uint32_t)(cd >> 32)) << 32) |
        (uint64_t)uniform_value<uint32_t>((uint32_t)cd);
  }
  u32x4 rnd_next = philox4x32_10({(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)c, TAG_METROPOLIS},
                                 sa.k0, sa.k1);
  int site_next = (int)bounded(uniform_value<uint32_t>(rnd_next.x), (uint32_t)N);
  float wn[JL];
#pragma unroll
  for (int jj = 0; jj < JL; ++jj) {
    const int j = lane + NKA_WAVE * jj;
    wn[jj] = j < M ? v.W[(size_t)site_next * M + j] : 0.0f;
  }
  const int n_sweeps = sa.n_discard + sa.chain_length;
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int p = 0; p < sa.sweep_size; ++p, ++q) {
      const u32x4 rnd = rnd_next;
      const int site = site_next;
      float w[JL];
#pragma unroll
      for (int jj = 0; jj < JL; ++jj) w[jj] = wn[jj];
      {
        const uint64_t qn = q + 1;
        rnd_next = philox4x32_10({(uint32_t)qn, (uint32_t)(qn >> 32), (uint32_t)c, TAG_METROPOLIS},
                                 sa.k0, sa.k1);
        site_next = (int)bounded(uniform_value<uint32_t>(rnd_next.x), (uint32_t)N);
#pragma unroll
        for (int jj = 0; jj < JL; ++jj) {
          const int j = lane + NKA_WAVE * jj;
          wn[jj] = j < M ? v.W[(size_t)site_next * M + j] : 0.0f;
        }
      }
      const bool up = xs[site] > 0;
      const float g = up ? -2.0f * kKappaS : 2.0f * kKappaS;
      float A = 0.0f, P = 1.0f;
#pragma unroll
      for (int jj = 0; jj < JL; ++jj) {
        const float t = fmaf(g, w[jj], th[jj]);
        A += fabsf(t) - fabsf(th[jj]);
        P = fmaf(P, __builtin_amdgcn_exp2f(-fabsf(t)), P);
      }
      float delta = wave_sum(A * (1.0f / kKappaS) + kLn2S * __builtin_amdgcn_logf(P * rP0));
      if (v.a) delta += up ? -2.0f * v.a[site] : 2.0f * v.a[site];
      delta = uniform_value<float>(delta);
      const float u = uniform_value<float>(uniform01<float>(rnd));
      if (u < exp_t<float>(pw * delta)) {
        // recompute t rather than keep it live through the reduction (an
        // opaque copy of g stops the compiler from reusing the values)
        float gr = g;
        asm volatile("" : "+v"(gr));
#pragma unroll
        for (int jj = 0; jj < JL; ++jj) th[jj] = fmaf(gr, w[jj], th[jj]);
        rP0 = 1.0f / P;
        if (lane == 0) xs[site] = (int8_t)(-xs[site]);
        lp += pw * delta;
        ++nacc;
        wave_lds_sync();
      }
    }
    if (sw >= sa.n_discard) {
      const int t = sw - sa.n_discard;
      int8_t* dst = samples + ((int64_t)c * sa.chain_length + t) * N;
      for (int i = lane; i < N; i += NKA_WAVE) dst[i] = xs[i];
      if (lp_samples && lane == 0) lp_samples[(int64_t)c * sa.chain_length + t] = lp;
    }
  }
  for (int i = lane; i < N; i += NKA_WAVE) sigma[(int64_t)c * N + i] = xs[i];
  if (lane == 0) {
    log_prob[c] = lp;
    n_accepted[c] += nacc;
  }
}

template <int JL>
__global__ void __launch_bounds__(256) rbm_metropolis_wide_kernel(nka_rbm m, SweepArgs sa, int8_t* __restrict__ sigma, float* __restrict__ log_prob, int8_t* __restrict__ samples, float* __restrict__ lp_samples, int64_t* __restrict__ n_accepted) {
  rbm_metropolis_wide_kernel_body<JL>(m, sa, sigma, log_prob, samples, lp_samples, n_accepted, blockIdx.x);
}

// The same sweeps for a batch of jobs of one shape: blockIdx.y selects the job
// (its parameters, chains, outputs and random stream); the shape arguments are shared.
template <int JL>
__global__ void __launch_bounds__(256) rbm_metropolis_wide_kernel_batched(const nka_sweep_job* __restrict__ jobs, SweepArgs sa) {
  const nka_sweep_job& j = jobs[blockIdx.y];
  SweepArgs s = sa;
  s.k0 = (uint32_t)j.seed;
  s.k1 = (uint32_t)(j.seed >> 32);
  s.counter = j.counter;
  s.counter_dev = j.counter_dev;
  rbm_metropolis_wide_kernel_body<JL>(j.model, s, j.sigma, (float*)j.log_prob, j.samples, (float*)j.log_prob_samples, j.n_accepted, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Several chains per wavefront.  For small RBMs (M <= 8 * 64/G hidden units)
// a proposal's arithmetic is a few hundred instructions that the per-proposal
// overhead (random draw, cluster bookkeeping, reduction, acceptance) would
// dominate on a wavefront per chain.  Here the wavefront holds G chains, one
// per group of LPC = 64/G lanes: lane gl of group g owns hidden units
// j = gl + LPC*jj, the spins of each chain are in LDS, and all per-chain
// quantities (the draw, the selected move, the sums) are equal across the
// group's lanes, so the groups run the same instruction stream on different
// chains.  Sums over a group use DPP row permutations whose partner lanes
// add the same two operands, so every lane of a group holds the identical sum.
// Same random streams, rules and acceptance as rbm_metropolis_kernel.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ inline float dpp_perm_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double dpp_perm_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL, class T> __device__ inline T dpp_perm(T v);
template <> __device__ inline float dpp_perm<0xb1, float>(float v) { return dpp_perm_f32<0xb1>(v); }
template <> __device__ inline float dpp_perm<0x4e, float>(float v) { return dpp_perm_f32<0x4e>(v); }
template <> __device__ inline float dpp_perm<0x141, float>(float v) { return dpp_perm_f32<0x141>(v); }
template <> __device__ inline float dpp_perm<0x140, float>(float v) { return dpp_perm_f32<0x140>(v); }
template <> __device__ inline double dpp_perm<0xb1, double>(double v) { return dpp_perm_f64<0xb1>(v); }
template <> __device__ inline double dpp_perm<0x4e, double>(double v) { return dpp_perm_f64<0x4e>(v); }
template <> __device__ inline double dpp_perm<0x141, double>(double v) { return dpp_perm_f64<0x141>(v); }
template <> __device__ inline double dpp_perm<0x140, double>(double v) { return dpp_perm_f64<0x140>(v); }

// Sum over the LPC lanes of this lane's group (LPC = 16 or 32), in every lane.
// log of a small positive integer count: hardware log2 for FP32, log for FP64.
template <class T> __device__ inline T log_fast(T x);
template <> __device__ inline float log_fast<float>(float x) {
  return 0.69314718055994531f * __builtin_amdgcn_logf(x);
}
template <> __device__ inline double log_fast<double>(double x) { return log(x); }

template <int LPC, class T>
__device__ inline T group_sum(T v) {
  v += dpp_perm<0xb1, T>(v);    // quad_perm [1,0,3,2]
  v += dpp_perm<0x4e, T>(v);    // quad_perm [2,3,0,1]
  v += dpp_perm<0x141, T>(v);   // row_half_mirror: quads 0 <-> 1 of each half-row
  v += dpp_perm<0x140, T>(v);   // row_mirror: half-rows 0 <-> 1 of each row
  if (LPC == 32) v += __shfl_xor(v, 16, NKA_WAVE);
  return v;
}

template <class T, int G, int JG, int RSG, bool STAGE>
__device__ __forceinline__ void rbm_metropolis_group_kernel_body(const nka_rbm& m, const SweepArgs& sa, int8_t* __restrict__ sigma, T* __restrict__ log_prob, int8_t* __restrict__ samples, T* __restrict__ lp_samples, int64_t* __restrict__ n_accepted, int bx) {
  constexpr int LPC = NKA_WAVE / G;
  constexpr uint64_t GMASK = (LPC == 64) ? ~0ull : ((1ull << LPC) - 1ull);
  extern __shared__ unsigned char smem[];
  const RbmView<T> v = rbm_view<T>(m);
  const int N = v.N, M = v.M;
  const int wid = threadIdx.x / NKA_WAVE, waves = blockDim.x / NKA_WAVE;
  const int lane = lane_id(), g = lane / LPC, gl = lane % LPC;
  const size_t wbytes = STAGE ? pad16((size_t)N * M * sizeof(T)) : 0;
  // LDS: [W (staged)] [visible bias a] [cluster pairs] [spins of each chain]
  T* a_s = (T*)(smem + wbytes);
  int32_t* cl_s = (int32_t*)(smem + wbytes + pad16((size_t)N * sizeof(T)));
  int8_t* xs = (int8_t*)(smem + wbytes + pad16((size_t)N * sizeof(T)) +
                         pad16((size_t)2 * sa.n_clusters * sizeof(int32_t))) +
               (wid * G + g) * pad16(N);
  const T* W = v.W;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a_s[i] = v.a ? v.a[i] : (T)0;
  for (int i = threadIdx.x; i < 2 * sa.n_clusters; i += blockDim.x) cl_s[i] = sa.clusters[i];
  if (STAGE) {
    stage_w<T>((T*)smem, v.W, N, M);  // ends with a workgroup barrier
    W = (const T*)smem;
  } else {
    __syncthreads();
  }
  const int c_raw = (bx * waves + wid) * G + g;
  if ((bx * waves + wid) * G >= sa.C) return;  // whole wavefront past the last chain
  const bool live = c_raw < sa.C;                        // groups past the last chain repeat it
  const int c = live ? c_raw : sa.C - 1;

  for (int i = gl; i < N; i += LPC) xs[i] = sigma[(int64_t)c * N + i];
  wave_lds_sync();
  // register slot jj of a lane holds hidden unit gl + LPC * ((jj + g) mod JG):
  // the rotation by the group index makes the groups of a wavefront read
  // different LDS bank ranges of their W rows in the same instruction
  T th[JG], lc[JG];
  T part = (T)0;
#pragma unroll
  for (int jj = 0; jj < JG; ++jj) {
    const int j = gl + LPC * ((jj + g) & (JG - 1));
    th[jj] = (T)0;
    lc[jj] = (T)0;
    if (j < M) {
      T acc = v.b ? v.b[j] : (T)0;
      for (int i = 0; i < N; ++i) {
        const T w = W[(size_t)i * M + j];
        acc += (xs[i] > 0) ? w : -w;
      }
      th[jj] = acc;
      lc[jj] = log_cosh<T>(acc);
      part += lc[jj];
    }
  }
  for (int i = gl; i < N; i += LPC) part += (xs[i] > 0) ? a_s[i] : -a_s[i];
  const T pw = (T)sa.machine_pow;
  T lp = pw * group_sum<LPC, T>(part);

  // cluster slots of this lane: cl = gl + LPC * r, r < RSG (RSG * LPC >= n_clusters)
  const bool exch = sa.rule == NKA_RULE_EXCHANGE;
  int cu[RSG], cv[RSG];
#pragma unroll
  for (int r = 0; r < RSG; ++r) {
    const int cl = gl + LPC * r;
    if (exch && cl < sa.n_clusters) {
      cu[r] = cl_s[2 * cl];
      cv[r] = cl_s[2 * cl + 1];
    } else {
      cu[r] = -1;
      cv[r] = -1;
    }
  }

  int64_t nacc = 0;
  uint64_t q = sa.counter;
  if (sa.counter_dev) {
    const uint64_t cd = *sa.counter_dev;
    q = ((uint64_t)uniform_value<uint32_t>((uint32_t)(cd >> 32)) << 32) |
        (uint64_t)uniform_value<uint32_t>((uint32_t)cd);
  }
  const int n_sweeps = sa.n_discard + sa.chain_length;
  u32x4 rnd_next = philox4x32_10({(uint32_t)q, (uint32_t)(q >> 32), (uint32_t)c, TAG_METROPOLIS},
                                 sa.k0, sa.k1);
  for (int sw = 0; sw < n_sweeps; ++sw) {
    for (int p = 0; p < sa.sweep_size; ++p, ++q) {
      const u32x4 rnd = rnd_next;
      {
        const uint64_t qn = q + 1;
        rnd_next = philox4x32_10({(uint32_t)qn, (uint32_t)(qn >> 32), (uint32_t)c, TAG_METROPOLIS},
                                 sa.k0, sa.k1);
      }
      int s0, s1 = 0, nf;
      T corr = (T)0;
      bool valid = true;
      if (!exch) {
        s0 = (int)bounded(rnd.x, (uint32_t)N);
        nf = 1;
      } else {
        // differing clusters of this group's chain as a bit set in cluster
        // order: cluster cl = gl + LPC * r is bit cl % 64 of word cl / 64
        constexpr int NW = (RSG * LPC + 63) / 64;
        uint64_t word[NW];
        int8_t xu[RSG], xv[RSG];
#pragma unroll
        for (int w = 0; w < NW; ++w) word[w] = 0;
#pragma unroll
        for (int r = 0; r < RSG; ++r) {
          xu[r] = cu[r] >= 0 ? xs[cu[r]] : (int8_t)0;
          xv[r] = cu[r] >= 0 ? xs[cv[r]] : (int8_t)0;
          const uint64_t gbits = (__ballot(xu[r] != xv[r]) >> (g * LPC)) & GMASK;
          word[(LPC * r) / 64] |= gbits << ((LPC * r) % 64);
        }
        int n_diff = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) n_diff += __popcll(word[w]);
        valid = n_diff > 0;
        int k = (int)bounded(rnd.x, (uint32_t)max(n_diff, 1));
        int sel = 0;
        bool found = false;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          const int cnt = __popcll(word[w]);
          if (!found && k < cnt) {
            sel = 64 * w + kth_set_bit(word[w], k);
            found = true;
          } else if (!found) {
            k -= cnt;
          }
        }
        if (!valid || sel >= sa.n_clusters) sel = 0;
        s0 = cl_s[2 * sel];
        s1 = cl_s[2 * sel + 1];
        nf = 2;
        // differing clusters after the swap (sigma'_u = sigma_v, sigma'_v = sigma_u)
        const int8_t y0 = xs[s0], y1 = xs[s1];
        int n_new = 0;
#pragma unroll
        for (int r = 0; r < RSG; ++r) {
          const int a = cu[r], bb = cv[r];
          const int8_t xa = a == s0 ? y1 : (a == s1 ? y0 : xu[r]);
          const int8_t xb = bb == s0 ? y1 : (bb == s1 ? y0 : xv[r]);
          n_new += __popcll((__ballot(a >= 0 && xa != xb) >> (g * LPC)) & GMASK);
        }
        corr = valid ? log_fast<T>((T)n_diff) - log_fast<T>((T)max(n_new, 1)) : (T)0;
      }
      // change of log psi with the new theta / log_cosh kept for a commit
      const int8_t x0 = xs[s0], x1 = xs[s1];
      T nth[JG], nlc[JG];
      T dpart = (T)0;
#pragma unroll
      for (int jj = 0; jj < JG; ++jj) {
        const int j = gl + LPC * ((jj + g) & (JG - 1));
        nth[jj] = th[jj];
        nlc[jj] = lc[jj];
        if (j < M) {
          const T w0 = W[(size_t)s0 * M + j];
          T d = (x0 > 0) ? -2 * w0 : 2 * w0;
          if (nf == 2) {
            const T w1 = W[(size_t)s1 * M + j];
            d += (x1 > 0) ? -2 * w1 : 2 * w1;
          }
          const T t = th[jj] + d;
          const T l = log_cosh<T>(t);
          nth[jj] = t;
          nlc[jj] = l;
          dpart += l - lc[jj];
        }
      }
      T delta = group_sum<LPC, T>(dpart);
      delta += (x0 > 0) ? -2 * a_s[s0] : 2 * a_s[s0];
      if (nf == 2) delta += (x1 > 0) ? -2 * a_s[s1] : 2 * a_s[s1];
      const T u = uniform01<T>(rnd);
      const bool acc = valid && (u < exp_t<T>(pw * delta + corr));
      if (acc) {
#pragma unroll
        for (int jj = 0; jj < JG; ++jj) {
          th[jj] = nth[jj];
          lc[jj] = nlc[jj];
        }
        if (gl == 0) {
          xs[s0] = (int8_t)(-x0);
          if (nf == 2) xs[s1] = (int8_t)(-x1);
        }
        lp += pw * delta;
        ++nacc;
      }
      wave_lds_sync();
    }
    if (sw >= sa.n_discard && live) {
      const int t = sw - sa.n_discard;
      int8_t* dst = samples + ((int64_t)c * sa.chain_length + t) * N;
      for (int i = gl; i < N; i += LPC) dst[i] = xs[i];
      if (lp_samples && gl == 0) lp_samples[(int64_t)c * sa.chain_length + t] = lp;
    }
  }
  if (live) {
    for (int i = gl; i < N; i += LPC) sigma[(int64_t)c * N + i] = xs[i];
    if (gl == 0) {
      log_prob[c] = lp;
      n_accepted[c] += nacc;
    }
  }
}

template <class T, int G, int JG, int RSG, bool STAGE>
__global__ void __launch_bounds__(256) rbm_metropolis_group_kernel(nka_rbm m, SweepArgs sa, int8_t* __restrict__ sigma, T* __restrict__ log_prob, int8_t* __restrict__ samples, T* __restrict__ lp_samples, int64_t* __restrict__ n_accepted) {
  rbm_metropolis_group_kernel_body<T, G, JG, RSG, STAGE>(m, sa, sigma, log_prob, samples, lp_samples, n_accepted, blockIdx.x);
}

// The same sweeps for a batch of jobs of one shape: blockIdx.y selects the job
// (its parameters, chains, outputs and random stream); the shape arguments are shared.
template <class T, int G, int JG, int RSG, bool STAGE>
__global__ void __launch_bounds__(256) rbm_metropolis_group_kernel_batched(const nka_sweep_job* __restrict__ jobs, SweepArgs sa) {
  const nka_sweep_job& j = jobs[blockIdx.y];
  SweepArgs s = sa;
  s.k0 = (uint32_t)j.seed;
  s.k1 = (uint32_t)(j.seed >> 32);
  s.counter = j.counter;
  s.counter_dev = j.counter_dev;
  rbm_metropolis_group_kernel_body<T, G, JG, RSG, STAGE>(j.model, s, j.sigma, (T*)j.log_prob, j.samples, (T*)j.log_prob_samples, j.n_accepted, blockIdx.x);
}

// counter += inc (one thread), after the sweep that consumed the counter
__global__ void advance_counter_kernel(uint64_t* __restrict__ c, uint64_t inc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = c[0] + inc;
}

// compute units of the current device (queried once per process)
static int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      cus = n;
    else
      cus = 256;
  }
  return cus;
}

template <class T>
static bool stage_fits_s(const nka_rbm& m, int waves) {
  return pad16((size_t)m.n_visible * m.n_hidden * sizeof(T)) + waves * pad16(m.n_visible) <=
         kLdsBudget;
}

template <class T>
static size_t group_lds(const nka_rbm& m, const SweepArgs& sa, int waves, int G) {
  return pad16((size_t)m.n_visible * m.n_hidden * sizeof(T)) + pad16((size_t)m.n_visible * sizeof(T)) +
         pad16((size_t)2 * sa.n_clusters * sizeof(int32_t)) + (size_t)waves * G * pad16(m.n_visible);
}

// Chains per wavefront for the grouped kernel (0: one wavefront per chain):
// 4 chains of 16 lanes when M <= 128 and there are chains enough for two such
// wavefronts per SIMD, else 2 of 32 lanes when M <= 256, provided the clusters
// fit the per-lane slots and W is staged in LDS.
template <class T>
static int pick_group(const nka_rbm& m, const SweepArgs& sa, int waves) {
  // four chains per wavefront only when that still leaves two wavefronts per
  // SIMD (the chains are latency-bound): measured on MI355X with 4096 chains,
  // M = 128: two chains per wavefront 0.79 ms, four 1.05 ms
  const bool four_ok = (int64_t)sa.C >= (int64_t)4 * 8 * device_cus();
  if (sa.weights) return 0;  // weighted exchange: one chain per wavefront
  for (int G : {4, 2}) {
    if (G == 4 && !four_ok) continue;
    const int lpc = NKA_WAVE / G;
    if (m.n_hidden > 8 * lpc) continue;
    if (sa.rule == NKA_RULE_EXCHANGE && sa.n_clusters > kMaxClusterSlots * lpc) continue;
    if (group_lds<T>(m, sa, waves, G) > kLdsBudget) continue;
    return G;
  }
  return 0;
}

// Every launch helper takes an optional batch: with jobs != nullptr the
// kernel runs once per job of the batch (grid y), each job's parameters,
// chains and random stream read from jobs[] (device memory), the shape from m.
template <class T, int G, int JG, int RSG>
static void launch_group_rs(const nka_rbm& m, const SweepArgs& sa, int8_t* sigma, T* lp, int8_t* smp,
                            T* lps, int64_t* nacc, int waves, const nka_sweep_job* jobs, int n_jobs,
                            hipStream_t s) {
  const size_t lds = group_lds<T>(m, sa, waves, G);
  const int per_block = waves * G;
  const unsigned g = (unsigned)((sa.C + per_block - 1) / per_block);
  if (jobs)
    hipLaunchKernelGGL((rbm_metropolis_group_kernel_batched<T, G, JG, RSG, true>), dim3(g, n_jobs),
                       dim3(64 * waves), lds, s, jobs, sa);
  else
    hipLaunchKernelGGL((rbm_metropolis_group_kernel<T, G, JG, RSG, true>), dim3(g), dim3(64 * waves), lds, s,
                       m, sa, sigma, lp, smp, lps, nacc);
}

// cluster slots per lane rounded up to a power of two (1 for the local rule)
template <class T, int G, int JG>
static void launch_group(const nka_rbm& m, const SweepArgs& sa, int8_t* sigma, T* lp, int8_t* smp,
                         T* lps, int64_t* nacc, int waves, const nka_sweep_job* jobs, int n_jobs,
                         hipStream_t s) {
  const int lpc = NKA_WAVE / G;
  const int need = sa.rule == NKA_RULE_EXCHANGE ? (sa.n_clusters + lpc - 1) / lpc : 1;
  if (need <= 1) launch_group_rs<T, G, JG, 1>(m, sa, sigma, lp, smp, lps, nacc, waves, jobs, n_jobs, s);
  else if (need <= 2) launch_group_rs<T, G, JG, 2>(m, sa, sigma, lp, smp, lps, nacc, waves, jobs, n_jobs, s);
  else if (need <= 4) launch_group_rs<T, G, JG, 4>(m, sa, sigma, lp, smp, lps, nacc, waves, jobs, n_jobs, s);
  else if (need <= 8) launch_group_rs<T, G, JG, 8>(m, sa, sigma, lp, smp, lps, nacc, waves, jobs, n_jobs, s);
  else launch_group_rs<T, G, JG, 16>(m, sa, sigma, lp, smp, lps, nacc, waves, jobs, n_jobs, s);
}

template <int JL, class T>
static void launch_sweep(const nka_rbm& m, const SweepArgs& sa, int8_t* sigma, T* lp, int8_t* smp,
                         T* lps, int64_t* nacc, const nka_sweep_job* jobs, int n_jobs, hipStream_t s) {
  const int waves = 4;
  const int G = pick_group<T>(m, sa, waves);
  if (G) {
    const int lpc = NKA_WAVE / G;
    const int need = (m.n_hidden + lpc - 1) / lpc;
    const int jg = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 8;
#define NKA_GROUP_CASE(GG, JJ)                                                                  \
    if (G == GG && jg == JJ) {                                                                  \
      launch_group<T, GG, JJ>(m, sa, sigma, lp, smp, lps, nacc, waves, jobs, n_jobs, s);        \
      return;                                                                                   \
    }
    NKA_GROUP_CASE(4, 1) NKA_GROUP_CASE(4, 2) NKA_GROUP_CASE(4, 4) NKA_GROUP_CASE(4, 8)
    NKA_GROUP_CASE(2, 1) NKA_GROUP_CASE(2, 2) NKA_GROUP_CASE(2, 4) NKA_GROUP_CASE(2, 8)
#undef NKA_GROUP_CASE
  }
  const bool stage = stage_fits_s<T>(m, waves);
  const size_t lds = (stage ? pad16((size_t)m.n_visible * m.n_hidden * sizeof(T)) : 0) +
                     waves * pad16(m.n_visible);
  const unsigned g = (unsigned)((sa.C + waves - 1) / waves);
  if constexpr (sizeof(T) == 4) {
    if (!stage && sa.rule == NKA_RULE_LOCAL) {
      // units per lane rounded up to a multiple of 8 (registers: theta' and
      // two W rows per unit)
      const int jlw = (m.n_hidden + 8 * NKA_WAVE - 1) / (8 * NKA_WAVE) * 8;
#define NKA_WIDE_CASE(J)                                                                                 \
      if (jlw == J) {                                                                                    \
        if (jobs)                                                                                        \
          hipLaunchKernelGGL((rbm_metropolis_wide_kernel_batched<J>), dim3(g, n_jobs), dim3(64 * waves), \
                             lds, s, jobs, sa);                                                          \
        else                                                                                             \
          hipLaunchKernelGGL((rbm_metropolis_wide_kernel<J>), dim3(g), dim3(64 * waves), lds, s, m, sa,  \
                             sigma, (float*)lp, smp, (float*)lps, nacc);                                 \
        return;                                                                                          \
      }
      NKA_WIDE_CASE(8) NKA_WIDE_CASE(16) NKA_WIDE_CASE(24) NKA_WIDE_CASE(32)
      NKA_WIDE_CASE(40) NKA_WIDE_CASE(48) NKA_WIDE_CASE(56) NKA_WIDE_CASE(64)
#undef NKA_WIDE_CASE
    }
  }
  if (jobs) {
    if (stage)
      hipLaunchKernelGGL((rbm_metropolis_kernel_batched<T, JL, true>), dim3(g, n_jobs), dim3(64 * waves), lds,
                         s, jobs, sa);
    else
      hipLaunchKernelGGL((rbm_metropolis_kernel_batched<T, JL, false>), dim3(g, n_jobs), dim3(64 * waves), lds,
                         s, jobs, sa);
    return;
  }
  if (stage)
    hipLaunchKernelGGL((rbm_metropolis_kernel<T, JL, true>), dim3(g), dim3(64 * waves), lds, s, m,
                       sa, sigma, lp, smp, lps, nacc);
  else
    hipLaunchKernelGGL((rbm_metropolis_kernel<T, JL, false>), dim3(g), dim3(64 * waves), lds, s,
                       m, sa, sigma, lp, smp, lps, nacc);
}
template <int JL>
static void launch_sweep_f32(const nka_rbm& m, const SweepArgs& sa, int8_t* sg, void* lp,
                             int8_t* smp, void* lps, int64_t* na, const nka_sweep_job* jobs, int n_jobs,
                             hipStream_t s) {
  launch_sweep<JL, float>(m, sa, sg, (float*)lp, smp, (float*)lps, na, jobs, n_jobs, s);
}
template <int JL>
static void launch_sweep_f64(const nka_rbm& m, const SweepArgs& sa, int8_t* sg, void* lp,
                             int8_t* smp, void* lps, int64_t* na, const nka_sweep_job* jobs, int n_jobs,
                             hipStream_t s) {
  launch_sweep<JL, double>(m, sa, sg, (double*)lp, smp, (double*)lps, na, jobs, n_jobs, s);
}

// counter += inc for every job of a batch that keeps its counter on the device
__global__ void advance_counters_kernel(const nka_sweep_job* __restrict__ jobs, int n_jobs, uint64_t inc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_jobs && jobs[k].counter_dev) jobs[k].counter_dev[0] = jobs[k].counter_dev[0] + inc;
}

}  // namespace nka

using namespace nka;

extern "C" {

int nka_random_states(int8_t* sigma, int n_chains, int n_sites, int n_down, uint64_t seed,
                      uint64_t counter, void* stream) {
  NKA_CHECK_ARG(n_chains >= 0 && n_sites > 0 && (n_chains == 0 || sigma),
                "nka_random_states: bad sizes");
  NKA_CHECK_ARG(n_down <= n_sites, "nka_random_states: n_down > n_sites");
  if (n_chains == 0) return NKA_OK;
  hipLaunchKernelGGL(random_states_kernel, dim3((n_chains + 127) / 128), dim3(128), 0,
                     (hipStream_t)stream, sigma, n_chains, n_sites, n_down, (uint32_t)seed,
                     (uint32_t)(seed >> 32), counter);
  NKA_CHECK_LAUNCH("nka_random_states");
  return NKA_OK;
}

int nka_rbm_metropolis(const nka_rbm* m, int rule, const int32_t* clusters, int n_clusters,
                       const uint32_t* cluster_weights, int8_t* sigma, void* log_prob, int n_chains, int sweep_size,
                       int n_discard, int chain_length, double machine_pow, uint64_t seed,
                       uint64_t counter, uint64_t* counter_dev, int8_t* samples,
                       void* log_prob_samples, int64_t* n_accepted, void* stream) {
  NKA_CHECK_ARG(m && m->params, "nka_rbm_metropolis: null model");
  NKA_CHECK_ARG(m->n_visible > 0 && m->n_hidden > 0 && m->n_hidden <= kMaxJL * NKA_WAVE,
                "nka_rbm_metropolis: unsupported sizes");
  NKA_CHECK_ARG(m->dtype == NKA_F32 || m->dtype == NKA_F64, "nka_rbm_metropolis: bad dtype");
  NKA_CHECK_ARG(rule == NKA_RULE_LOCAL || rule == NKA_RULE_EXCHANGE,
                "nka_rbm_metropolis: unknown rule %d", rule);
  NKA_CHECK_ARG(rule == NKA_RULE_LOCAL ||
                    (clusters && n_clusters > 0 && n_clusters <= kMaxClusterSlots * NKA_WAVE),
                "nka_rbm_metropolis: exchange rule needs 1..%d clusters",
                kMaxClusterSlots * NKA_WAVE);
  NKA_CHECK_ARG(n_chains >= 0 && sweep_size >= 0 && n_discard >= 0 && chain_length >= 0,
                "nka_rbm_metropolis: bad sizes");
  NKA_CHECK_ARG(n_chains == 0 || (sigma && log_prob && n_accepted &&
                                  (chain_length == 0 || samples)),
                "nka_rbm_metropolis: null buffer");
  if (n_chains == 0) return NKA_OK;
  SweepArgs sa;
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
  const int jl = pick_jl(m->n_hidden);
  hipStream_t s = (hipStream_t)stream;
  if (m->dtype == NKA_F32) {
    NKA_DISPATCH_JL(jl, launch_sweep_f32, *m, sa, sigma, log_prob, samples, log_prob_samples,
                    n_accepted, (const nka_sweep_job*)nullptr, 0, s);
  } else {
    NKA_DISPATCH_JL(jl, launch_sweep_f64, *m, sa, sigma, log_prob, samples, log_prob_samples,
                    n_accepted, (const nka_sweep_job*)nullptr, 0, s);
  }
  if (counter_dev) {
    const uint64_t inc = (uint64_t)(n_discard + chain_length) * (uint64_t)sweep_size;
    hipLaunchKernelGGL(advance_counter_kernel, dim3(1), dim3(64), 0, s, counter_dev, inc);
  }
  NKA_CHECK_LAUNCH("nka_rbm_metropolis");
  return NKA_OK;
}

int nka_rbm_metropolis_batched(const nka_sweep_job* jobs, int n_jobs, const nka_rbm* shape, int rule,
                               const int32_t* clusters, int n_clusters, int n_chains, int sweep_size,
                               int n_discard, int chain_length, double machine_pow, void* stream) {
  NKA_CHECK_ARG(shape && jobs && n_jobs > 0 && n_jobs <= 65535, "nka_rbm_metropolis_batched: bad batch");
  NKA_CHECK_ARG(shape->n_visible > 0 && shape->n_hidden > 0 && shape->n_hidden <= kMaxJL * NKA_WAVE,
                "nka_rbm_metropolis_batched: unsupported sizes");
  NKA_CHECK_ARG(shape->dtype == NKA_F32 || shape->dtype == NKA_F64, "nka_rbm_metropolis_batched: bad dtype");
  NKA_CHECK_ARG(rule == NKA_RULE_LOCAL || rule == NKA_RULE_EXCHANGE,
                "nka_rbm_metropolis_batched: unknown rule %d", rule);
  NKA_CHECK_ARG(rule == NKA_RULE_LOCAL ||
                    (clusters && n_clusters > 0 && n_clusters <= kMaxClusterSlots * NKA_WAVE),
                "nka_rbm_metropolis_batched: exchange rule needs 1..%d clusters",
                kMaxClusterSlots * NKA_WAVE);
  NKA_CHECK_ARG(n_chains > 0 && sweep_size >= 0 && n_discard >= 0 && chain_length >= 0,
                "nka_rbm_metropolis_batched: bad sizes");
  SweepArgs sa;
  sa.rule = rule;
  sa.clusters = clusters;
  sa.n_clusters = rule == NKA_RULE_EXCHANGE ? n_clusters : 0;
  sa.weights = nullptr;
  sa.C = n_chains;
  sa.sweep_size = sweep_size;
  sa.n_discard = n_discard;
  sa.chain_length = chain_length;
  sa.machine_pow = machine_pow;
  sa.k0 = sa.k1 = 0;
  sa.counter = 0;
  sa.counter_dev = nullptr;
  const int jl = pick_jl(shape->n_hidden);
  hipStream_t s = (hipStream_t)stream;
  if (shape->dtype == NKA_F32) {
    NKA_DISPATCH_JL(jl, launch_sweep_f32, *shape, sa, (int8_t*)nullptr, (void*)nullptr, (int8_t*)nullptr,
                    (void*)nullptr, (int64_t*)nullptr, jobs, n_jobs, s);
  } else {
    NKA_DISPATCH_JL(jl, launch_sweep_f64, *shape, sa, (int8_t*)nullptr, (void*)nullptr, (int8_t*)nullptr,
                    (void*)nullptr, (int64_t*)nullptr, jobs, n_jobs, s);
  }
  const uint64_t inc = (uint64_t)(n_discard + chain_length) * (uint64_t)sweep_size;
  hipLaunchKernelGGL(advance_counters_kernel, dim3((n_jobs + 255) / 256), dim3(256), 0, s, jobs, n_jobs, inc);
  NKA_CHECK_LAUNCH("nka_rbm_metropolis_batched");
  return NKA_OK;
}

}  // extern "C"
