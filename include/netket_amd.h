/*
 * netket_amd — C ABI of the MI355X (gfx950) VMC step.
 *
 * Every entry point takes caller-owned device buffers, plain sizes and a HIP
 * stream (passed as void*), enqueues its kernels on that stream and returns
 * 0 on success or a negative status (nka_last_error() holds the message).
 * Nothing here allocates device memory or synchronises the stream; workspace
 * buffers are passed in and their sizes are queryable.
 *
 * Element types are selected by the dtype codes NKA_F32 / NKA_F64.  Spin
 * configurations are int8 arrays of +1/-1 (netket/hilbert/spin.py:165-178).
 *
 * The entry points replace the JAX functions of NetKet's VMC step; each
 * declaration cites the reference function (file:line) it stands in for.
 */
#ifndef NETKET_AMD_H
#define NETKET_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NKA_F32 0
#define NKA_F64 1

#define NKA_OK 0
#define NKA_ERR_ARG -1
#define NKA_ERR_HIP -2
#define NKA_ERR_UNSUPPORTED -3

/* Library identification and the message of the last failed call (per thread). */
const char* nka_version(void);
const char* nka_last_error(void);

/* ------------------------------------------------------------------------
 * Operators: connected elements (DiscreteJaxOperator.get_conn_padded,
 * netket/operator/_discrete_operator_jax.py:117-139).
 * ---------------------------------------------------------------------- */

/* IsingJax.get_conn_padded (netket/operator/_ising/jax.py:82-88, 125-165).
 * x: (n_batch, n_sites) int8; edges: (n_edges, 2) int32;
 * xp: (n_batch, n_rows, n_sites) int8; mels: (n_batch, n_rows) of mel_dtype.
 * n_rows is n_sites + 1, or 1 when h == 0 (the StaticZero branch). */
int nka_ising_conn(const int8_t* x, int64_t n_batch, int n_sites, const int32_t* edges,
                   int n_edges, double h, double J, int mel_dtype, int n_rows, int8_t* xp,
                   void* mels, void* stream);

/* Packed lookup tables of a spin-1/2 LocalOperator, the device image of
 * pack_internals_jax (netket/operator/_local_operator/compile_helpers.py:29-218).
 * Terms are ordered by acting size, then insertion order.  For term t with
 * ns = op_nsites[t] sites op_sites[4t..4t+ns) (sorted), the row of a
 * configuration is r = sum_k idx(x[site_k]) * 2^(ns-1-k) with idx(+1)=0,
 * idx(-1)=1, and its table entry is row = op_row_off[t] + r.  row_diag[row] is
 * the diagonal element; the off-diagonal elements with |mel| > mel_cutoff are
 * conn_mel[row_conn_off[row] + k], k < row_nconn[row], with new local indices
 * conn_newidx[...] (same bit order as r) and owning term conn_term[...].
 * All pointers are device pointers. */
typedef struct nka_localop {
  int n_ops;
  int max_conn_size;
  int nonzero_diagonal;
  int mel_dtype;
  double constant;
  double mel_cutoff;
  const int32_t* op_sites;
  const int32_t* op_nsites;
  const int32_t* op_row_off;
  const void* row_diag;
  const int32_t* row_nconn;
  const int32_t* row_conn_off;
  const void* conn_mel;
  const int32_t* conn_newidx;
  const int32_t* conn_term;
} nka_localop;

/* LocalOperatorJax._get_conn_padded (netket/operator/_local_operator/jax.py:74-201,
 * 256-288): xp (n_batch, max_conn_size, n_sites), mels (n_batch, max_conn_size),
 * n_conn (n_batch,) may be NULL. */
int nka_localop_conn(const int8_t* x, int64_t n_batch, int n_sites, const nka_localop* op,
                     int8_t* xp, void* mels, int32_t* n_conn, void* stream);

/* ------------------------------------------------------------------------
 * RBM ansatz (netket/models/rbm.py:57-81, netket/nn/activation.py:78-84).
 * Parameters live in one flat device buffer in the reference's ravel order
 * (jax pytree key order): Dense/bias (M), Dense/kernel (N x M, row-major),
 * visible_bias (N).  An absent bias has offset -1.
 * ---------------------------------------------------------------------- */
typedef struct nka_rbm {
  int n_visible;
  int n_hidden;
  int dtype;
  int reserved;
  const void* params;
  int64_t bias_off;
  int64_t kernel_off;
  int64_t visible_off;
} nka_rbm;

/* RBM.__call__: log psi of n configurations (netket/models/rbm.py:57-81). */
int nka_rbm_log_psi(const nka_rbm* m, const int8_t* sigma, int64_t n, void* log_psi,
                    void* stream);

/* Operator selector for the fused local-energy kernel. */
#define NKA_OP_ISING 0
#define NKA_OP_LOCAL 1
typedef struct nka_ising {
  const int32_t* edges;
  int n_edges;
  int mel_dtype;
  double h;
  double J;
} nka_ising;

/* local_value_kernel_jax (netket/vqs/mc/kernels.py:62-71) fused with
 * get_conn_padded: eloc[s] = sum_k mel_k exp(log psi(x'_k) - log psi(x_s)),
 * evaluated with the hidden-unit update theta' = theta - 2 sigma_i W_i. so the
 * connected configurations are never materialised.  eloc is FP64; log_psi
 * (model dtype) may be NULL.  An Ising operator with a transverse field on an
 * FP32 model whose W exceeds 64 KiB takes a path tiled over samples and hidden
 * units that needs a workspace: nka_rbm_local_energy_workspace returns its
 * size for n samples (0 when no workspace is needed); a smaller one, of at
 * least the size for 64 samples, processes the batch in chunks. */
int64_t nka_rbm_local_energy_workspace(const nka_rbm* m, int op_kind, const void* op, int64_t n);
int nka_rbm_local_energy(const nka_rbm* m, int op_kind, const void* op, const int8_t* sigma,
                         int64_t n, double* eloc, void* log_psi, void* workspace,
                         int64_t workspace_bytes, void* stream);

/* The local energies of a batch of jobs of one shape, Ising operators on one
 * lattice (each job its own h and J), in one launch (the serving loop's shared
 * launches).  jobs is an array in device memory; shape and op_shape give the
 * shared sizes and edges (their parameter pointer and couplings are not read);
 * n samples per job.  Each job's values equal nka_rbm_local_energy's.  Models
 * that take the tiled path are refused (nka_rbm_local_energy_batchable
 * returns 0 for them). */
typedef struct nka_eloc_job {
  nka_rbm model;
  nka_ising op;
  const int8_t* sigma;
  double* eloc;
  void* log_psi;
} nka_eloc_job;
int nka_rbm_local_energy_batchable(const nka_rbm* shape, const nka_ising* op);
int nka_rbm_local_energy_batched(const nka_eloc_job* jobs, int n_jobs, const nka_rbm* shape,
                                 const nka_ising* op_shape, int64_t n, void* stream);

/* ------------------------------------------------------------------------
 * Two-layer dense network: log psi(x) = sum_k log_cosh(z2_k) with
 * z2 = log_cosh(x W1 + b1) W2 + b2 (Dense -> log_cosh -> Dense -> log_cosh ->
 * sum: nn.Dense of netket/nn/blocks/mlp.py:105-115, log_cosh of
 * netket/nn/activation.py:78-84; the user module of the J1-J2 workload).
 * Flat parameters in ravel order Dense_0/bias (H1), Dense_0/kernel (N, H1),
 * Dense_1/bias (H2), Dense_1/kernel (H1, H2); a bias offset is -1 when the
 * layer has no bias.  n_hidden1, n_hidden2 <= 1024.
 * ---------------------------------------------------------------------- */
typedef struct nka_ffnn {
  int n_visible;
  int n_hidden1;
  int n_hidden2;
  int dtype;
  const void* params;
  int64_t b1_off;
  int64_t w1_off;
  int64_t b2_off;
  int64_t w2_off;
} nka_ffnn;

/* log psi of n configurations (model dtype). */
int nka_ffnn_log_psi(const nka_ffnn* m, const int8_t* sigma, int64_t n, void* log_psi, void* stream);

/* local_value_kernel_jax (netket/vqs/mc/kernels.py:62-71) with
 * get_conn_padded of an Ising or LocalOperator (op_kind as for
 * nka_rbm_local_energy): eloc[s] = sum_k mel_k exp(log psi(x'_k) - log psi(x_s))
 * (FP64).  The connected configurations are formed in chunks in the
 * workspace (nka_ffnn_local_energy_workspace bytes suffice; a smaller
 * workspace gives smaller chunks); the first layer of x'_k is updated from
 * x_s by the sites where they differ.  log_psi (model dtype) may be NULL. */
int64_t nka_ffnn_local_energy_workspace(const nka_ffnn* m, int op_kind, const void* op, int64_t n);
int nka_ffnn_local_energy(const nka_ffnn* m, int op_kind, const void* op, const int8_t* sigma, int64_t n,
                          double* eloc, void* log_psi, void* workspace, int64_t workspace_bytes,
                          void* stream);

/* Metropolis sweeps for the network, with the arguments and semantics of
 * nka_rbm_metropolis (netket/sampler/metropolis.py:382-505). */
int nka_ffnn_metropolis(const nka_ffnn* m, int rule, const int32_t* clusters, int n_clusters,
                        const uint32_t* cluster_weights, int8_t* sigma, void* log_prob, int n_chains, int sweep_size, int n_discard,
                        int chain_length, double machine_pow, uint64_t seed, uint64_t counter,
                        uint64_t* counter_dev, int8_t* samples, void* log_prob_samples,
                        int64_t* n_accepted, void* stream);

/* Per-sample factors of the log-derivatives (model dtype): h1 = log_cosh(theta1)
 * (ns, H1), d1 = d log psi / d theta1 = (W2 tanh(z2)) * tanh(theta1) (ns, H1),
 * g2 = tanh(z2) (ns, H2); the Jacobian is O = [d1 | x (x) d1 | g2 | h1 (x) g2]
 * in ravel order (nkjax.jacobian, netket/jax/_jacobian/logic.py:308-387).
 * workspace holds nka_ffnn_prepare_workspace(...) bytes. */
int64_t nka_ffnn_prepare_workspace(const nka_ffnn* m, int64_t ns);
int nka_ffnn_prepare(const nka_ffnn* m, const int8_t* sigma, int64_t ns, void* h1, void* d1, void* g2,
                     void* workspace, void* stream);

/* out[v, p] = sum_s (O[s, p] - mean[p]) w[v, s] (FP64, n_w <= 2), as
 * nka_rbm_contract, from the factors of nka_ffnn_prepare. */
int64_t nka_ffnn_contract_workspace(const nka_ffnn* m, int64_t ns, int n_w);
int nka_ffnn_contract(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1,
                      const void* g2, int64_t ns, const double* w, int n_w, const double* mean,
                      double* out, void* workspace, void* stream);

/* Centred NTK without the Jacobian, as nka_rbm_ntk:
 * (O O^T)_st = (hb2 + h1_s.h1_t)(g2_s.g2_t) + (hb1 + x_s.x_t)(d1_s.d1_t), centred
 * with v = O mean and c = |mean|^2; out = scale * O_c O_c^T + shift I + offset 11^T
 * (lower triangle, FP64; netket/_src/ngd/srt.py:55-66 with scale = 1/ns). */
int64_t nka_ffnn_ntk_workspace(const nka_ffnn* m, int64_t ns);
int nka_ffnn_ntk(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1, const void* g2,
                 int64_t ns, const double* mean, double scale, double shift, double offset, double* out,
                 int64_t ldo, void* workspace, void* stream);

/* Jacobian-vector product of the centred Jacobian, as nka_rbm_jvp, from the
 * factors of nka_ffnn_prepare (netket/_src/ngd/srt.py:34-35). */
int64_t nka_ffnn_jvp_workspace(const nka_ffnn* m, int64_t ns);
int nka_ffnn_jvp(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1, const void* g2,
                 int64_t ns, const double* v, const double* mean, double alpha, double beta, double* y,
                 void* workspace, void* stream);

/* Centred, rescaled Jacobian (layouts and output kinds of nka_rbm_jacobian). */
int nka_ffnn_jacobian(const nka_ffnn* m, const int8_t* sigma, const void* h1, const void* d1,
                      const void* g2, int64_t ns, const double* mean, double scale, int layout,
                      int out_kind, void* out_a, void* out_b, int64_t ldo, void* stream);

/* ------------------------------------------------------------------------
 * Sampler.
 * ---------------------------------------------------------------------- */

/* Random initial chain states (MetropolisSampler._init_state,
 * netket/sampler/metropolis.py:353-380): uniform +-1 when n_down < 0, else a
 * uniformly random placement of n_down spins -1 (the SumConstraint branch,
 * netket/hilbert/random/homogeneous.py:50-72, netket/hilbert/random/fock.py:77-97). */
int nka_random_states(int8_t* sigma, int n_chains, int n_sites, int n_down, uint64_t seed,
                      uint64_t counter, void* stream);

#define NKA_RULE_LOCAL 0
#define NKA_RULE_EXCHANGE 1

/* MetropolisSampler reset + sample(n_discard) + sample(chain_length)
 * (netket/vqs/mc/mc_state/state.py:521-576; netket/sampler/metropolis.py:382-505)
 * with LocalRule (netket/sampler/rules/local.py:40-49) or ExchangeRule
 * (netket/sampler/rules/exchange.py:143-218) for an RBM.  One sweep is
 * sweep_size proposals; after each of the last chain_length sweeps the chain
 * state is written to samples (n_chains, chain_length, n_sites).  sigma
 * (n_chains, n_sites) and log_prob (n_chains,, model dtype) are the persistent
 * chain state.  Proposal q of chain c draws Philox4x32-10 at counter
 * (counter + q, c, "METR") with key seed.  When counter_dev is not NULL the
 * base counter is read from that device word instead and advanced there by
 * (n_discard + chain_length) * sweep_size after the sweep, so a captured
 * graph replays with fresh draws.  clusters: (n_clusters, 2) int32 (exchange
 * only).  cluster_weights (exchange only, may be NULL for equal weights):
 * (n_clusters,) integers in [1, 2^20], device memory; a differing cluster is
 * proposed with probability weight / (total weight of the differing
 * clusters) and the log-correction is log total - log total' (ExchangeRule's
 * `probabilities`, netket/sampler/rules/exchange.py:86-90, 147-184).
 * n_accepted (n_chains,) is incremented. */
int nka_rbm_metropolis(const nka_rbm* m, int rule, const int32_t* clusters, int n_clusters,
                       const uint32_t* cluster_weights, int8_t* sigma, void* log_prob, int n_chains, int sweep_size,
                       int n_discard, int chain_length, double machine_pow, uint64_t seed,
                       uint64_t counter, uint64_t* counter_dev, int8_t* samples,
                       void* log_prob_samples, int64_t* n_accepted, void* stream);

/* The sweeps of a batch of independent jobs of one shape in one launch (the
 * serving loop's shared launches; NetKet runs one sampler per process, so
 * this has no single counterpart beyond nka_rbm_metropolis above).  Job k's
 * parameters, chains, outputs and Philox stream are in jobs[k] (an array in
 * device memory); shape gives n_visible, n_hidden, dtype and the parameter
 * offsets shared by all jobs (its params pointer is not read); the other
 * arguments are shared.  Each job's chains are exactly those that
 * nka_rbm_metropolis gives it alone. */
typedef struct nka_sweep_job {
  nka_rbm model;
  int8_t* sigma;
  void* log_prob;
  int8_t* samples;
  void* log_prob_samples;
  int64_t* n_accepted;
  uint64_t seed;
  uint64_t counter;
  uint64_t* counter_dev;
} nka_sweep_job;
int nka_rbm_metropolis_batched(const nka_sweep_job* jobs, int n_jobs, const nka_rbm* shape, int rule,
                               const int32_t* clusters, int n_clusters, int n_chains, int sweep_size,
                               int n_discard, int chain_length, double machine_pow, void* stream);

/* ------------------------------------------------------------------------
 * Statistics (netket/stats/mc_stats_old.py:87-196): out[0..7] =
 * {mean, var, var(chain means), n_blocks, var(block means), var(half-chain means),
 *  block_size, n} of data (n_chains, chain_length) in FP64 (variances with ddof 0).
 * ---------------------------------------------------------------------- */
int nka_stats_moments(const double* data, int n_chains, int chain_length, double* out8,
                      void* stream);

/* ------------------------------------------------------------------------
 * Jacobian of the RBM (nkjax.jacobian, netket/jax/_jacobian/logic.py:308-387)
 * and the quantities built from it.
 * ---------------------------------------------------------------------- */

/* t = tanh(theta), theta = sigma W + b: t (ns, M) and tT (M, ns) in model dtype. */
int nka_rbm_prepare(const nka_rbm* m, const int8_t* sigma, int64_t ns, void* t, void* tT,
                    void* stream);

/* Weighted sums over samples of the per-sample log-derivatives:
 * out[v, p] = sum_s (O[s, p] - mean[p]) * w[v, s] for v < n_w <= 2 (FP64),
 * O[s,p] = d log psi(sigma_s) / d theta_p in the ravel order, t = tanh(theta)
 * (ns, M) from nka_rbm_prepare, and mean == NULL meaning no centring.  With
 * w = 1/ns this is the column mean used for centring (logic.py:365-369); with
 * w = 2 (E - mean E)/ns it is the gradient 2 Re F
 * (netket/vqs/mc/mc_state/expect_forces.py:98-104, netket/vqs/mc/common.py:102-118);
 * with centring and w = x / sqrt(ns) it is O_L^T x (netket/_src/ngd/srt.py:79,
 * netket/_src/ngd/sr.py:43).  workspace holds nka_rbm_contract_workspace(...) bytes. */
int64_t nka_rbm_contract_workspace(const nka_rbm* m, int64_t ns, int n_w);
int nka_rbm_contract(const nka_rbm* m, const int8_t* sigma, const void* t, int64_t ns,
                     const double* w, int n_w, const double* mean, double* out, void* workspace,
                     void* stream);

/* Centred, rescaled Jacobian O_L[s,p] = (O[s,p] - mean[p]) * scale
 * (logic.py:365-373, netket/_src/ngd/sr_srt_common.py:75-80).
 * layout NKA_LAYOUT_PS writes X = O_L^T (n_params x ns), NKA_LAYOUT_SP writes
 * X = O_L (ns x n_params); both row-major.  kind NKA_OUT_BF16X2 writes the
 * bf16 split X = hi + lo into out_a (hi) and out_b (lo); NKA_OUT_F32 /
 * NKA_OUT_F64 write out_a only.  Rows of the output are ldo elements apart;
 * columns past the row length are not touched. */
#define NKA_LAYOUT_PS 0
#define NKA_LAYOUT_SP 1
#define NKA_OUT_BF16X2 0
#define NKA_OUT_F32 1
#define NKA_OUT_F64 2
int nka_rbm_jacobian(const nka_rbm* m, const int8_t* sigma, const void* t, const void* tT,
                     int64_t ns, const double* mean, double scale, int layout, int out_kind,
                     void* out_a, void* out_b, int64_t ldo, void* stream);

/* ------------------------------------------------------------------------
 * QGT / NTK Gram matrices and dense FP64 linear algebra.
 * ---------------------------------------------------------------------- */

/* out = X X^T + shift I + offset 11^T (lower triangle incl. diagonal, FP64,
 * row-major ldo) from the bf16 split X = hi + lo (n x k, row-major ldx) with
 * three BF16 MFMA products (hi hi^T + hi lo^T + lo hi^T) accumulated in FP32.
 * S = O_L^T O_L (QGTJacobianDenseT.to_dense,
 * netket/optimizer/qgt/qgt_jacobian_dense.py:136-161; netket/_src/ngd/sr.py:58-63)
 * or T = O_L O_L^T (netket/_src/ngd/srt.py:55-66). */
int nka_gram_bf16x3(const uint16_t* x_hi, const uint16_t* x_lo, int64_t n, int64_t k,
                    int64_t ldx, double shift, double offset, double* out, int64_t ldo,
                    void* stream);

/* Dense S of the RBM without forming its Jacobian (replaces the
 * matrix = O_L^T O_L + diag_shift I of netket/_src/ngd/sr.py:43-55 and
 * QGTJacobianDenseT.to_dense, netket/optimizer/qgt/qgt_jacobian_dense.py:29-161):
 * out = scale * O_c^T O_c + shift I (lower triangle incl. diagonal, FP64,
 * row-major ldo >= n_params), O_c[s,p] = O[s,p] - mean_s O[s,p], O as for
 * nka_rbm_contract, tT the (n_hidden x ns) activations of nka_rbm_prepare.
 * Each O_p is a product (augmented spin) x (augmented tanh), so
 * sum_s O_sp O_sq is one entry of C = P Q^T over the samples, P the distinct
 * spin-pair products (exact in bf16) and Q the distinct pair products of the
 * centred activations (bf16 hi + lo): a third of the matrix-core work of the
 * bf16x3 Gram of the Jacobian.  The column sums of O come from the same
 * product, so the centring cancels the large terms exactly.  FP32 models
 * with +-1 spins only; workspace holds nka_rbm_qgt_workspace(m, ns) bytes. */
int64_t nka_rbm_qgt_workspace(const nka_rbm* m, int64_t ns);
int nka_rbm_qgt(const nka_rbm* m, const int8_t* sigma, const void* tT, int64_t ns, double scale, double shift,
                double* out, int64_t ldo, void* workspace, void* stream);

/* Centred NTK of the RBM without forming its Jacobian: out = scale * O_c O_c^T
 * + shift I + offset 11^T (lower triangle incl. diagonal, FP64, row-major ldo),
 * O_c[s,p] = O[s,p] - mean[p], O as for nka_rbm_contract and t from
 * nka_rbm_prepare.  It uses (O O^T)_st = (hb + sigma_s.sigma_t)(t_s.t_t)
 * + ha sigma_s.sigma_t (hb, ha = 1 with a hidden / visible bias), two FP64
 * products with inner dimension n_hidden and n_visible instead of n_params.
 * With scale = 1/ns this is T = O_L O_L^T of netket/_src/ngd/srt.py:55-66.
 * FP64 models use FP64 matrix cores; FP32 models form t.t on bf16 matrix
 * cores from t = hi + lo (hi hi^T + hi lo^T + lo hi^T, FP32 accumulation) and
 * sigma.sigma exactly.  workspace holds nka_rbm_ntk_workspace(...) bytes. */
int64_t nka_rbm_ntk_workspace(const nka_rbm* m, int64_t ns);
int nka_rbm_ntk(const nka_rbm* m, const int8_t* sigma, const void* t, int64_t ns,
                const double* mean, double scale, double shift, double offset, double* out,
                int64_t ldo, void* workspace, void* stream);

/* Jacobian-vector product of the centred Jacobian without forming it:
 * y = alpha (O - 1 mean^T) v + beta y (FP64, ns entries; beta == 0: y is not
 * read), O as for nka_rbm_contract, t from nka_rbm_prepare, v of n_params
 * entries in ravel order, mean == NULL meaning no centring.  With
 * alpha = -mu / sqrt(ns), beta = 1 and y = dv this is the SPRING momentum term
 * dv -= mu O_L old_updates of netket/_src/ngd/srt.py:34-35.  workspace holds
 * nka_rbm_jvp_workspace(...) bytes. */
int64_t nka_rbm_jvp_workspace(const nka_rbm* m, int64_t ns);
int nka_rbm_jvp(const nka_rbm* m, const int8_t* sigma, const void* t, int64_t ns, const double* v,
                const double* mean, double alpha, double beta, double* y, void* workspace, void* stream);

/* C = beta C + alpha X X^T on the lower triangle (FP64 MFMA), X n x k row-major. */
int nka_syrk_f64(const double* x, int64_t n, int64_t k, int64_t ldx, double alpha, double beta,
                 double shift, double* c, int64_t ldc, void* stream);

/* C = beta C + alpha A B^T (FP64 MFMA); A m x k, B n x k, C m x n, row-major. */
int nka_gemm_nt_f64(const double* a, int64_t lda, const double* b, int64_t ldb, double* c,
                    int64_t ldc, int64_t m, int64_t n, int64_t k, double alpha, double beta,
                    void* stream);

/* y = alpha A x + beta y (trans = 0) or alpha A^T x + beta y (trans = 1); A m x n
 * row-major FP64, or the lower triangle of a symmetric matrix when trans = 2. */
int nka_gemv_f64(const double* a, int64_t lda, int64_t m, int64_t n, int trans, const double* x,
                 double alpha, double beta, double* y, void* workspace, void* stream);

/* Cholesky factorisation A = L L^T of the lower triangle in place (FP64),
 * cho_factor of netket/optimizer/solver/solvers.py:241-270.  workspace holds
 * nka_potrf_workspace(n) bytes (inverses of the diagonal blocks).  info (one
 * device int32) is 0 on success, or j+1 for the first non-positive or
 * non-finite pivot j; later kernels of the same call skip their work then. */
int64_t nka_potrf_workspace(int64_t n);
int nka_potrf_f64(double* a, int64_t n, int64_t lda, void* workspace, int32_t* info,
                  void* stream);

/* cho_solve: b <- A^{-1} b from the factor of nka_potrf_f64 (same workspace);
 * b is set to NaN when the factorisation failed (info != 0). */
int nka_potrs_f64(const double* a, int64_t n, int64_t lda, const void* workspace, double* b,
                  const int32_t* info, void* stream);

/* Cholesky factorisation and solve in one call with no allocation and no
 * caller workspace: a = L L^T in place (lower triangle) and b <- A^{-1} b, as
 * nka_potrf_f64 followed by nka_potrs_f64.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Its scratch is a static
 * device array of fixed size, one per GPU, held until the process ends:
 * n <= nka_cholesky_solve_noalloc_max_n() (32768, about 17 MB); larger n is
 * rejected with NKA_ERR_ARG.  All calls on one GPU share that scratch, so
 * calls that overlap on one GPU (different streams, no ordering) race.
 * info is cleared by the call, then set as by nka_potrf_f64; b becomes NaN
 * when the factorisation fails.  Results equal those of the workspace pair up
 * to the summation order of the split tail tiles of the trailing products. */
int64_t nka_cholesky_solve_noalloc_max_n(void);
int nka_cholesky_solve_f64_noalloc(double* a, int64_t n, int64_t lda, double* b, int32_t* info,
                                   void* stream);

/* x <- src (or 0 when src is NULL) if *flag != 0, else unchanged: the
 * device-side branch of the nan_fallback solver
 * (netket/_src/solvers/nan_fallback.py:34-38, jax.lax.cond on the flag). */
int nka_select_on_flag_f64(const int32_t* flag, const double* src, double* x, int64_t n, void* stream);

/* Scale-invariant regularisation of the dense QGT, diag_scale of
 * netket/optimizer/qgt/qgt_jacobian_common.py:21-60 (to_shift_offset, rescale)
 * and qgt_jacobian_dense.py:93-150: scale_k = sqrt(S_kk + offset) from the
 * unregularised S in the lower triangle of a; mode 0 replaces the lower
 * triangle by S_ij / (scale_i scale_j) + shift delta_ij (the system that is
 * solved), mode 1 adds shift scale_k^2 to the diagonal (to_dense). */
int nka_qgt_rescale_f64(double* a, int64_t n, int64_t lda, double offset, double shift, int mode,
                        double* scale, void* stream);
/* out = x * d (divide = 0) or x / d (divide = 1), FP64 elementwise: the y / scale
 * and x / scale of the rescaled solve (qgt_jacobian_dense.py:99-132). */
int nka_mul_div_f64(const double* x, const double* d, int64_t n, int divide, double* out, void* stream);

/* Mirror the lower triangle of a row-major FP64 matrix onto its upper triangle. */
int nka_symmetrize_f64(double* a, int64_t n, int64_t lda, void* stream);

/* Flags non-finite entries: *flag |= 1 if any of x[0..n) is NaN/Inf (the
 * nan_fallback test of netket/_src/solvers/nan_fallback.py:22-45). */
int nka_check_finite(const double* x, int64_t n, int32_t* flag, void* stream);

/* Conjugate gradient for A x = b (A full row-major FP64 n x n, e.g. S + eps I),
 * with the semantics of jax.scipy.sparse.linalg.cg, the default solver of
 * nk.optimizer.SR (netket/optimizer/sr.py:121): x0 = 0; iterate while
 * r.r > max(tol^2 b.b, atol^2) and k < maxiter.  nka_cg_start initialises x
 * and the workspace (nka_cg_workspace(n) bytes); each nka_cg_iterate call
 * enqueues `iters` iterations that become no-ops once converged;
 * nka_cg_status copies {r.r, p.Ap, alpha, beta, atol^2, k, done, maxiter} into
 * a device buffer of 8 doubles. */
int64_t nka_cg_workspace(int64_t n);
int nka_cg_start(const double* b, int64_t n, double tol, double atol, int64_t maxiter, double* x,
                 void* workspace, void* stream);
/* As nka_cg_start from a given x0 (r0 = b - A x0; the tolerance stays
 * relative to b.b): the x0 that SR passes when solver_restart=True
 * (netket/optimizer/preconditioner.py:118). */
int nka_cg_start_x0(const double* a, int64_t lda, const double* b, int64_t n, double tol, double atol,
                    int64_t maxiter, const double* x0, double* x, void* workspace, void* stream);
int nka_cg_iterate(const double* a, int64_t lda, int64_t n, double* x, void* workspace, int iters,
                   void* stream);
int nka_cg_status(const void* workspace, int64_t n, double* out8, void* stream);

/* ------------------------------------------------------------------------
 * Parameter update: optax.sgd via apply_gradient
 * (netket/_src/driver/abstract_optimization_driver.py:86-92,
 * netket/optimizer/__init__.py:33-60): p <- p + (-lr * (dtype)dp).
 * ---------------------------------------------------------------------- */
int nka_sgd_update(void* params, int dtype, const double* dp, int64_t n, double lr,
                   void* stream);

/* Elementwise helpers used by the host-side solvers (no reference counterpart:
 * building blocks of the CG / dv preparation). */
int nka_centre_scale(const double* e, int64_t n, const double* mean, double scale, double* out,
                     void* stream);
/* y += alpha x (FP64): the momentum term updates += mu old_updates of
 * netket/_src/ngd/srt.py:80-82. */
int nka_axpy_f64(int64_t n, double alpha, const double* x, double* y, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NETKET_AMD_H */
