// Connected elements <x|H|x'> of the two operator families on the VMC path.
//
// nka_ising_conn   : IsingJax.get_conn_padded (netket/operator/_ising/jax.py:125-165)
// nka_localop_conn : LocalOperatorJax.get_conn_padded
//                    (netket/operator/_local_operator/jax.py:74-201)
//
// Both write the padded layout of the reference bit for bit: xp (B, K, N)
// int8 spin values and mels (B, K) in the operator dtype.
#include <stdarg.h>
#include <stdio.h>

#include "nka_common.hpp"
#include "netket_amd.h"

namespace nka {

static thread_local char g_err[512];
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }

// ---------------------------------------------------------------------------
// Ising.  Row 0: x, mel J * sum_e (2*[x_u == x_v] - 1) (an exact integer sum
// times J).  Row r >= 1: x with site r-1 flipped, mel -h.
// ---------------------------------------------------------------------------
__global__ void ising_xp_kernel(const int8_t* __restrict__ x, int64_t B, int N, int K,
                                int8_t* __restrict__ xp) {
  const int64_t total = B * (int64_t)K * N;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx % N);
    const int64_t br = idx / N;
    const int r = (int)(br % K);
    const int64_t b = br / K;
    const int8_t v = x[b * N + i];
    xp[idx] = (r >= 1 && i == r - 1) ? (int8_t)(-v) : v;
  }
}

template <class T>
__global__ void ising_mel_kernel(const int8_t* __restrict__ x, int64_t B, int N, int K,
                                 const int32_t* __restrict__ edges, int E, T h, T J,
                                 T* __restrict__ mels) {
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < B;
       b += (int64_t)gridDim.x * blockDim.x) {
    const int8_t* xb = x + b * N;
    int s = 0;
    for (int e = 0; e < E; ++e) s += (xb[edges[2 * e]] == xb[edges[2 * e + 1]]) ? 1 : -1;
    T* m = mels + b * K;
    m[0] = J * (T)s;
    for (int r = 1; r < K; ++r) m[r] = -h;
  }
}

// ---------------------------------------------------------------------------
// Local operators.  One wavefront per sample.
//
// Entry order of the reference (before compaction): [diagonal if the operator
// has a nonzero diagonal], then for each term (grouped by acting size, which
// the host already did) the off-diagonal elements of the row selected by x.
// Compaction keeps entries with |mel| > mel_cutoff in order and pads with
// (x, 0) up to max_conn_size (jax.py:189-199).  Every stored off-diagonal
// element already satisfies |mel| > cutoff, so only the diagonal can drop.
// ---------------------------------------------------------------------------
template <class T>
__device__ inline int term_row(const int8_t* xs, const int32_t* sites, int ns) {
  int r = 0;
  for (int k = 0; k < ns; ++k) r = 2 * r + (xs[sites[k]] < 0 ? 1 : 0);
  return r;
}

template <class T>
__global__ void localop_conn_kernel(const int8_t* __restrict__ x, int64_t B, int N,
                                    nka_localop op, int8_t* __restrict__ xp,
                                    T* __restrict__ mels, int32_t* __restrict__ n_conn_out) {
  extern __shared__ unsigned char smem[];
  const int wid = threadIdx.x / NKA_WAVE;
  const int lane = lane_id();
  const int waves = blockDim.x / NKA_WAVE;
  const int K = op.max_conn_size;
  // per-wave LDS: x (N bytes, padded to 16), entry table (K ints)
  const int xbytes = (N + 15) & ~15;
  unsigned char* base = smem + (size_t)wid * (xbytes + 4 * (size_t)K);
  int8_t* xs = (int8_t*)base;
  int32_t* entry = (int32_t*)(base + xbytes);  // encoded (term, k) or -1 = diag

  const int32_t* op_sites = op.op_sites;
  const T* row_diag = (const T*)op.row_diag;
  const T* conn_mel = (const T*)op.conn_mel;

  for (int64_t b = (int64_t)blockIdx.x * waves + wid; b < B; b += (int64_t)gridDim.x * waves) {
    for (int i = lane; i < N; i += NKA_WAVE) xs[i] = x[b * N + i];
    wave_lds_sync();

    // diagonal: constant + (sum over groups of the in-order sum of the group's
    // terms), jax.py:117-122
    int count = 0;
    if (lane == 0) {
      double d = 0.0, gsum = 0.0;
      int gsize = op.n_ops > 0 ? op.op_nsites[0] : 0;
      for (int t = 0; t < op.n_ops; ++t) {
        if (op.op_nsites[t] != gsize) {
          d += gsum;
          gsum = 0.0;
          gsize = op.op_nsites[t];
        }
        const int r = term_row<T>(xs, op_sites + 4 * t, op.op_nsites[t]);
        gsum += (double)row_diag[op.op_row_off[t] + r];
      }
      d += gsum;
      const T diag = (T)(op.constant + d);
      int e = 0;
      if (op.nonzero_diagonal && fabs((double)diag) > op.mel_cutoff) {
        entry[e++] = -1;
        mels[b * K] = diag;
      }
      // off-diagonal entries in term order
      for (int t = 0; t < op.n_ops; ++t) {
        const int r = term_row<T>(xs, op_sites + 4 * t, op.op_nsites[t]);
        const int row = op.op_row_off[t] + r;
        const int nc = op.row_nconn[row];
        const int off = op.row_conn_off[row];
        for (int k = 0; k < nc && e < K; ++k) {
          entry[e] = off + k;
          mels[b * K + e] = conn_mel[off + k];
          ++e;
        }
      }
      for (int p = e; p < K; ++p) {
        entry[p] = -2;
        mels[b * K + p] = (T)0;
      }
      if (n_conn_out) n_conn_out[b] = e;
      count = e;
    }
    count = __shfl(count, 0, NKA_WAVE);
    wave_lds_sync();

    // rows: x with the sites of the selected term set to the column's states
    for (int e = 0; e < K; ++e) {
      const int code = entry[e];
      int8_t* row = xp + (b * K + e) * (int64_t)N;
      int t = -1, newidx = 0, ns = 0;
      int s0 = -1, s1 = -1, s2 = -1, s3 = -1;
      if (code >= 0) {
        // find the term owning conn index `code` (binary search on row_conn_off is
        // avoided: the host stores the owning term's sites in op_sites via conn2term)
        t = op.conn_term[code];
        newidx = op.conn_newidx[code];
        ns = op.op_nsites[t];
        s0 = op_sites[4 * t + 0];
        s1 = op_sites[4 * t + 1];
        s2 = op_sites[4 * t + 2];
        s3 = op_sites[4 * t + 3];
      }
      for (int i = lane; i < N; i += NKA_WAVE) {
        int8_t v = xs[i];
        if (code >= 0) {
          int k = -1;
          if (i == s0) k = 0;
          else if (i == s1) k = 1;
          else if (i == s2) k = 2;
          else if (i == s3) k = 3;
          if (k >= 0 && k < ns) {
            const int li = (newidx >> (ns - 1 - k)) & 1;
            v = li ? (int8_t)-1 : (int8_t)1;
          }
        }
        row[i] = v;
      }
    }
    (void)count;
    wave_lds_sync();
  }
}

}  // namespace nka

using namespace nka;

extern "C" {

const char* nka_last_error(void) { return nka::last_error(); }

const char* nka_version(void) { return "netket_amd 0.1 (gfx950)"; }

int nka_ising_conn(const int8_t* x, int64_t n_batch, int n_sites, const int32_t* edges,
                   int n_edges, double h, double J, int mel_dtype, int n_rows, int8_t* xp,
                   void* mels, void* stream) {
  NKA_CHECK_ARG(n_batch >= 0 && n_sites > 0, "nka_ising_conn: bad sizes");
  NKA_CHECK_ARG(n_rows == 1 || n_rows == n_sites + 1, "nka_ising_conn: n_rows must be 1 or N+1");
  NKA_CHECK_ARG(n_edges >= 0 && (n_edges == 0 || edges), "nka_ising_conn: bad edges");
  NKA_CHECK_ARG(mel_dtype == NKA_F32 || mel_dtype == NKA_F64, "nka_ising_conn: bad dtype");
  if (n_batch == 0) return NKA_OK;
  hipStream_t s = (hipStream_t)stream;
  const int64_t total = n_batch * (int64_t)n_rows * n_sites;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(ising_xp_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, n_batch,
                     n_sites, n_rows, xp);
  int64_t mb = (n_batch + 255) / 256;
  if (mb > 65536) mb = 65536;
  if (mel_dtype == NKA_F64)
    hipLaunchKernelGGL(ising_mel_kernel<double>, dim3((unsigned)mb), dim3(256), 0, s, x,
                       n_batch, n_sites, n_rows, edges, n_edges, h, J, (double*)mels);
  else
    hipLaunchKernelGGL(ising_mel_kernel<float>, dim3((unsigned)mb), dim3(256), 0, s, x,
                       n_batch, n_sites, n_rows, edges, n_edges, (float)h, (float)J,
                       (float*)mels);
  NKA_CHECK_LAUNCH("nka_ising_conn");
  return NKA_OK;
}

int nka_localop_conn(const int8_t* x, int64_t n_batch, int n_sites, const nka_localop* op,
                     int8_t* xp, void* mels, int32_t* n_conn, void* stream) {
  NKA_CHECK_ARG(op != nullptr, "nka_localop_conn: null operator");
  NKA_CHECK_ARG(n_batch >= 0 && n_sites > 0, "nka_localop_conn: bad sizes");
  NKA_CHECK_ARG(op->max_conn_size >= 1, "nka_localop_conn: max_conn_size must be >= 1");
  NKA_CHECK_ARG(op->mel_dtype == NKA_F32 || op->mel_dtype == NKA_F64,
                "nka_localop_conn: bad dtype");
  if (n_batch == 0) return NKA_OK;
  hipStream_t s = (hipStream_t)stream;
  const int waves = 4;
  const size_t per_wave = (size_t)((n_sites + 15) & ~15) + 4 * (size_t)op->max_conn_size;
  const size_t lds = per_wave * waves;
  NKA_CHECK_ARG(lds <= 64 * 1024, "nka_localop_conn: N or max_conn_size too large (%zu B LDS)",
                lds);
  int64_t blocks = (n_batch + waves - 1) / waves;
  if (blocks > 16384) blocks = 16384;
  if (op->mel_dtype == NKA_F64)
    hipLaunchKernelGGL(localop_conn_kernel<double>, dim3((unsigned)blocks), dim3(64 * waves), lds,
                       s, x, n_batch, n_sites, *op, xp, (double*)mels, n_conn);
  else
    hipLaunchKernelGGL(localop_conn_kernel<float>, dim3((unsigned)blocks), dim3(64 * waves), lds,
                       s, x, n_batch, n_sites, *op, xp, (float*)mels, n_conn);
  NKA_CHECK_LAUNCH("nka_localop_conn");
  return NKA_OK;
}

}  // extern "C"
