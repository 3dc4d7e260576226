// Conjugate gradient on the device, with jax.scipy.sparse.linalg.cg semantics
// (the default SR solver, netket/optimizer/sr.py:121): x0 = 0 (or a given x0,
// used when SR keeps the previous solution, netket/optimizer/preconditioner.py:118), stop when
// r.r <= max(tol^2 b.b, atol^2) or after maxiter iterations.
//
// State lives in a small device array so that no iteration synchronises with
// the host:  sc[0] gamma = r.r, sc[1] p.Ap, sc[2] alpha, sc[3] beta,
// sc[4] atol^2, sc[5] iteration count, sc[6] done flag, sc[7] maxiter.
// Partial dot products use a fixed grid, so results are deterministic.
#include "nka_common.hpp"
#include "netket_amd.h"

namespace nka {

constexpr int CG_BLOCKS = 256;

__device__ inline double block_reduce(double v, double* red) {
  v = wave_sum(v);
  const int wid = threadIdx.x / 64;
  if (lane_id() == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x / 64); ++w) t += red[w];
  __syncthreads();
  return t;
}

// part[blockIdx] = sum over the block's slice of x.y
__global__ void __launch_bounds__(256) dot_partial(const double* __restrict__ x, const double* __restrict__ y,
                                                   int64_t n, double* __restrict__ part,
                                                   const double* __restrict__ sc) {
  __shared__ double red[4];
  if (sc && sc[6] != 0.0) return;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  s = block_reduce(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// sc[slot] = sum of part (one block)
__global__ void __launch_bounds__(256) dot_final(const double* __restrict__ part, int nparts,
                                                 double* __restrict__ sc, int slot, int skip_if_done) {
  __shared__ double red[4];
  if (skip_if_done && sc[6] != 0.0) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  s = block_reduce(s, red);
  if (threadIdx.x == 0) sc[slot] = s;
}

// init: r = b, p = b, x = 0
__global__ void cg_init_vec(const double* __restrict__ b, int64_t n, double* __restrict__ x,
                            double* __restrict__ r, double* __restrict__ p) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] = 0.0;
    r[i] = b[i];
    p[i] = b[i];
  }
}

// sc: gamma already = b.b in sc[0]; set atol2, k = 0, done flag, maxiter
__global__ void cg_init_scalars(double* __restrict__ sc, double tol, double atol, double maxiter) {
  const double bb = sc[0];
  const double a2 = fmax(tol * tol * bb, atol * atol);
  sc[4] = a2;
  sc[5] = 0.0;
  sc[7] = maxiter;
  sc[6] = (bb <= a2 || maxiter <= 0.0) ? 1.0 : 0.0;
}

// x += alpha p, r -= alpha Ap, with alpha = gamma / pAp; partial r.r
__global__ void __launch_bounds__(256) cg_update(double* __restrict__ x, double* __restrict__ r,
                                                 const double* __restrict__ p, const double* __restrict__ Ap,
                                                 int64_t n, const double* __restrict__ sc,
                                                 double* __restrict__ part) {
  __shared__ double red[4];
  if (sc[6] != 0.0) return;
  const double alpha = sc[0] / sc[1];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    s += ri * ri;
  }
  s = block_reduce(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// gamma_new from partials; beta; counters; done flag
__global__ void __launch_bounds__(256) cg_finalize(const double* __restrict__ part, int nparts,
                                                   double* __restrict__ sc) {
  __shared__ double red[4];
  if (sc[6] != 0.0) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  s = block_reduce(s, red);
  if (threadIdx.x == 0) {
    sc[3] = s / sc[0];
    sc[0] = s;
    sc[5] += 1.0;
    if (s <= sc[4] || sc[5] >= sc[7]) sc[6] = 2.0;  // 2: finished this iteration
  }
}

// p = r + beta p (also run in the iteration that finished, harmlessly), then
// promote the "finished this iteration" mark to done.
__global__ void cg_p_update(double* __restrict__ p, const double* __restrict__ r, int64_t n,
                            const double* __restrict__ sc) {
  if (sc[6] == 1.0) return;
  const double beta = sc[3];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = r[i] + beta * p[i];
}

__global__ void cg_mark_done(double* __restrict__ sc) {
  if (sc[6] == 2.0) sc[6] = 1.0;
}

// y = A x for a full row-major matrix, skipped once CG is done.
__global__ void __launch_bounds__(256) cg_matvec(const double* __restrict__ A, int64_t lda, int64_t n,
                                                 const double* __restrict__ x, double* __restrict__ y,
                                                 const double* __restrict__ sc) {
  if (sc[6] != 0.0) return;
  const int wid = threadIdx.x / 64, lane = lane_id();
  for (int64_t r = (int64_t)blockIdx.x * 4 + wid; r < n; r += (int64_t)gridDim.x * 4) {
    double s = 0.0;
    for (int64_t c = lane; c < n; c += 64) s += A[r * lda + c] * x[c];
    s = wave_sum(s);
    if (lane == 0) y[r] = s;
  }
}

// x = x0, r = b - A x0, p = r for a full row-major matrix (one wavefront per row).
__global__ void __launch_bounds__(256) cg_residual(const double* __restrict__ A, int64_t lda, int64_t n,
                                                   const double* __restrict__ b, const double* __restrict__ x0,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ p) {
  const int wid = threadIdx.x / 64, lane = lane_id();
  for (int64_t row = (int64_t)blockIdx.x * 4 + wid; row < n; row += (int64_t)gridDim.x * 4) {
    double s = 0.0;
    for (int64_t c = lane; c < n; c += 64) s += A[row * lda + c] * x0[c];
    s = wave_sum(s);
    if (lane == 0) {
      const double ri = b[row] - s;
      x[row] = x0[row];
      r[row] = ri;
      p[row] = ri;
    }
  }
}

// sc[0] = r0.r0 and sc[1] = b.b already; the tolerance is relative to b.b
// as in jax.scipy.sparse.linalg.cg, the first test is made on r0.r0.
__global__ void cg_init_scalars_x0(double* __restrict__ sc, double tol, double atol, double maxiter) {
  const double a2 = fmax(tol * tol * sc[1], atol * atol);
  sc[4] = a2;
  sc[5] = 0.0;
  sc[7] = maxiter;
  sc[6] = (sc[0] <= a2 || maxiter <= 0.0) ? 1.0 : 0.0;
}

}  // namespace nka

using namespace nka;

extern "C" {

int64_t nka_cg_workspace(int64_t n) { return (4 * n + CG_BLOCKS + 8) * (int64_t)sizeof(double); }

int nka_cg_start(const double* b, int64_t n, double tol, double atol, int64_t maxiter, double* x,
                 void* workspace, void* stream) {
  NKA_CHECK_ARG(n > 0 && b && x && workspace, "nka_cg_start: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  double* ws = (double*)workspace;
  double* r = ws;
  double* p = ws + n;
  double* part = ws + 3 * n;
  double* sc = ws + 3 * n + CG_BLOCKS;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, CG_BLOCKS);
  hipLaunchKernelGGL(cg_init_vec, dim3(g), dim3(256), 0, s, b, n, x, r, p);
  hipLaunchKernelGGL(dot_partial, dim3(CG_BLOCKS), dim3(256), 0, s, b, b, n, part, nullptr);
  hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, s, part, CG_BLOCKS, sc, 0, 0);
  hipLaunchKernelGGL(cg_init_scalars, dim3(1), dim3(1), 0, s, sc, tol, atol, (double)maxiter);
  NKA_CHECK_LAUNCH("nka_cg_start");
  return NKA_OK;
}

int nka_cg_start_x0(const double* a, int64_t lda, const double* b, int64_t n, double tol, double atol,
                    int64_t maxiter, const double* x0, double* x, void* workspace, void* stream) {
  NKA_CHECK_ARG(n > 0 && lda >= n && a && b && x0 && x && workspace, "nka_cg_start_x0: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  double* ws = (double*)workspace;
  double* r = ws;
  double* p = ws + n;
  double* part = ws + 3 * n;
  double* sc = ws + 3 * n + CG_BLOCKS;
  hipLaunchKernelGGL(dot_partial, dim3(CG_BLOCKS), dim3(256), 0, s, b, b, n, part, nullptr);
  hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, s, part, CG_BLOCKS, sc, 1, 0);
  const unsigned gm = (unsigned)std::min<int64_t>((n + 3) / 4, 4096);
  hipLaunchKernelGGL(cg_residual, dim3(gm), dim3(256), 0, s, a, lda, n, b, x0, x, r, p);
  hipLaunchKernelGGL(dot_partial, dim3(CG_BLOCKS), dim3(256), 0, s, r, r, n, part, nullptr);
  hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, s, part, CG_BLOCKS, sc, 0, 0);
  hipLaunchKernelGGL(cg_init_scalars_x0, dim3(1), dim3(1), 0, s, sc, tol, atol, (double)maxiter);
  NKA_CHECK_LAUNCH("nka_cg_start_x0");
  return NKA_OK;
}

int nka_cg_iterate(const double* A, int64_t lda, int64_t n, double* x, void* workspace, int iters,
                   void* stream) {
  NKA_CHECK_ARG(n > 0 && A && x && workspace && iters >= 0, "nka_cg_iterate: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  double* ws = (double*)workspace;
  double* r = ws;
  double* p = ws + n;
  double* Ap = ws + 2 * n;
  double* part = ws + 3 * n;
  double* sc = ws + 3 * n + CG_BLOCKS;
  const unsigned gm = (unsigned)std::min<int64_t>((n + 3) / 4, 4096);
  const unsigned gv = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(cg_matvec, dim3(gm), dim3(256), 0, s, A, lda, n, p, Ap, sc);
    hipLaunchKernelGGL(dot_partial, dim3(CG_BLOCKS), dim3(256), 0, s, p, Ap, n, part, sc);
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, s, part, CG_BLOCKS, sc, 1, 1);
    hipLaunchKernelGGL(cg_update, dim3(CG_BLOCKS), dim3(256), 0, s, x, r, p, Ap, n, sc, part);
    hipLaunchKernelGGL(cg_finalize, dim3(1), dim3(256), 0, s, part, CG_BLOCKS, sc);
    hipLaunchKernelGGL(cg_p_update, dim3(gv), dim3(256), 0, s, p, r, n, sc);
    hipLaunchKernelGGL(cg_mark_done, dim3(1), dim3(1), 0, s, sc);
  }
  NKA_CHECK_LAUNCH("nka_cg_iterate");
  return NKA_OK;
}

// Copy the 8 CG scalars (gamma, pAp, alpha, beta, atol2, k, done, maxiter) to out.
int nka_cg_status(const void* workspace, int64_t n, double* out8, void* stream) {
  NKA_CHECK_ARG(workspace && out8 && n > 0, "nka_cg_status: bad arguments");
  const double* sc = (const double*)workspace + 3 * n + CG_BLOCKS;
  hipError_t e = hipMemcpyAsync(out8, sc, 8 * sizeof(double), hipMemcpyDeviceToDevice,
                                (hipStream_t)stream);
  if (e != hipSuccess) {
    set_error("nka_cg_status: %s", hipGetErrorString(e));
    return NKA_ERR_HIP;
  }
  return NKA_OK;
}

}  // extern "C"
