/*
 * rbmhip.h -- C ABI of the MI355X (gfx950) variational-Monte-Carlo library
 * librbmhip.so.  Plain pointers and sizes only; every function returns
 * RBMHIP_OK (0) or a negative error code, and rbmhip_last_error() returns the
 * message of the calling thread's last failure.  No exception crosses the ABI.
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) doubles;
 *   - a spin configuration is two uint64 words, spin i = +1 iff bit i is set
 *     (reference include/Spins.h:125-129), N <= 128;
 *   - model, operator and ensemble handles are host objects; device copies are
 *     made on first use on the current device and refreshed after a setter;
 *   - every compute call runs on the library's stream of the current device and
 *     returns after the results are on the host, like the reference's blocking
 *     calls; calls from several host threads must not share a handle;
 *   - compute calls need a GPU and fail with RBMHIP_ERR_NO_DEVICE otherwise.
 *
 * Each declaration cites the reference interface it replaces (file:line,
 * relative to the reference repository root).
 */
#ifndef RBMHIP_H
#define RBMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBMHIP_OK 0
#define RBMHIP_ERR_ARG (-1)
#define RBMHIP_ERR_NO_DEVICE (-2)
#define RBMHIP_ERR_HIP (-3)
#define RBMHIP_ERR_UNSUPPORTED (-4)
#define RBMHIP_ERR_NUMERIC (-5)

#define RBMHIP_FP32 32 /* complex64 kernels, FP64 reductions */
#define RBMHIP_FP64 64 /* complex128 kernels */
#define RBMHIP_BF16 16 /* PsiDeep only: forward passes as BF16 MFMA layers with FP32
                          accumulation and activations; O_k by the complex64 kernels on
                          the bfloat16-rounded weights */

typedef struct rbmhip_psi rbmhip_psi;
typedef struct rbmhip_op rbmhip_op;
typedef struct rbmhip_ens rbmhip_ens;
typedef struct rbmhip_made rbmhip_made;

/* ---- library / device ---------------------------------------------------- */
const char* rbmhip_last_error(void);
int rbmhip_version(void);
/* setDevice: source/RBMonGPU.cu:7-9 (bound at pyRBMonGPU/main.cpp:611) */
int rbmhip_set_device(int device);
int rbmhip_device_count(int* count);
/* start_profiling / stop_profiling: source/RBMonGPU.cu:11-17 (main.cpp:612-613).
 * The reference's cudaProfilerStart/Stop become roctxProfilerResume/Pause
 * (collection control for rocprofv3 --selected-regions) around a roctx range
 * named "rbmhip profiled region"; both synchronise the library stream first. */
int rbmhip_start_profiling(void);
int rbmhip_stop_profiling(void);
/* activation_function (my_logcosh), pyRBMonGPU/main.cpp:607-609;
 * derivative = 1 evaluates my_tanh (include/quantum_state/psi_functions.hpp:97-114) */
int rbmhip_activation(const double* z, int64_t n, int derivative, double* out);

/* ---- wavefunctions ------------------------------------------------------- */
/* Psi(alpha, beta, b, W, prefactor, free_quantum_axis, gpu): include/quantum_state/Psi.hpp:222-239.
 * b: M complex, W: N x M complex row-major, a: optional visible bias (N complex,
 * an extension with no counterpart in the reference; NULL keeps the reference model). */
int rbmhip_rbm_create(int N, int M, const double* b, const double* W, const double* a, double prefactor,
                      int precision, rbmhip_psi** out);
/* Psi::set_params (source/quantum_state/Psi.cu:80-92), device side */
int rbmhip_rbm_set_params(rbmhip_psi* psi, const double* b, const double* W, const double* a);
/* PsiDeep(alpha, beta, biases, connections, W, final_weights, prefactor, ...):
 * include/quantum_state/PsiDeep.hpp:339-439.  Layers are flattened: sizes[L],
 * conn_counts[L] (C_l), conns = concatenated C_l x size_l uint32 tables,
 * weights = concatenated C_l x size_l complex, biases = concatenated size_l
 * complex, final_weights = size_{L-1} complex.  ref_compat_final = 1 keeps the
 * reference's final-weight derivative log_psi * x_j (PsiDeep.hpp:285-291);
 * 0 uses dlog psi/df_j = x_j.  precision = RBMHIP_BF16 needs layers of at most
 * 256 units without translational invariance, and weights that fit in LDS. */
int rbmhip_deep_create(int N, int num_layers, const int32_t* sizes, const int32_t* conn_counts,
                       const uint32_t* conns, const double* weights, const double* biases,
                       const double* final_weights, double prefactor, int translational_invariance,
                       int ref_compat_final, int precision, rbmhip_psi** out);
/* PsiDeepT::set_params: source/quantum_state/PsiDeep.cu:143-177 */
int rbmhip_deep_set_params(rbmhip_psi* psi, const double* weights, const double* biases, const double* final_weights);
/* PsiDeep.translational_invariance (pyRBMonGPU/main.cpp:110) */
int rbmhip_deep_set_translational_invariance(rbmhip_psi* psi, int on);
/* prefactor property (main.cpp:64, 109) */
int rbmhip_psi_set_prefactor(rbmhip_psi* psi, double prefactor);
/* num_params (main.cpp:89, 114): 2N + M + NM (+ N with visible bias) for the RBM,
 * source/quantum_state/PsiDeep.cu:36-63 for PsiDeep */
int rbmhip_psi_num_params(const rbmhip_psi* psi, int64_t* out);
int rbmhip_psi_destroy(rbmhip_psi* psi);

/* ---- operators ----------------------------------------------------------- */
/* Operator(coefficients, pauli_types, pauli_indices, gpu): include/operator/Operator.hpp:212-227.
 * types: S x L int32 (0 = I, 1 = X, 2 = Y, 3 = Z); idx: S x L int32, each
 * string ends at the first -1 (a string without one continues into the next
 * row, as in Operator.hpp:148-152). */
int rbmhip_op_create(int S, int L, const double* coefficients, const int32_t* types, const int32_t* indices,
                     rbmhip_op** out);
/* Canonical form: strings grouped by flip mask.  hermitian = 1 when the
 * operator's matrix equals its conjugate transpose. */
int rbmhip_op_info(const rbmhip_op* op, int* num_diagonal, int* num_flip_masks, int* hermitian, int* max_index);
int rbmhip_op_destroy(rbmhip_op* op);

/* ---- spin ensembles ------------------------------------------------------ */
/* ExactSummation(N, gpu): source/spin_ensembles/ExactSummation.cu:16-24 */
int rbmhip_exact_create(int N, rbmhip_ens** out);
/* ExactSummation::set_total_z_symmetry: ExactSummation.cu:26-45 */
int rbmhip_exact_set_total_z(rbmhip_ens* ens, int sector);
/* MonteCarloLoop(num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains, gpu):
 * source/spin_ensembles/MonteCarloLoop.cu:20-42.  seed keys the Philox stream
 * (the reference seeds curand with 0, MonteCarloLoop.cu:15). */
int rbmhip_mc_create(uint32_t num_samples, uint32_t num_sweeps, uint32_t num_thermalization_sweeps,
                     uint32_t num_markov_chains, uint64_t seed, rbmhip_ens** out);
/* MonteCarloLoop(MonteCarloLoop&): MonteCarloLoop.cu:44-66 */
int rbmhip_mc_copy(const rbmhip_ens* other, rbmhip_ens** out);
/* set_total_z_symmetry / set_fast_sweep: include/spin_ensembles/MonteCarloLoop.hpp:283-291 */
int rbmhip_mc_set_total_z(rbmhip_ens* ens, int sector);
int rbmhip_mc_set_fast_sweep(rbmhip_ens* ens, uint32_t num_tries);
/* Persistent chains (opt-in, no counterpart in the reference, which restarts
 * every chain from a random configuration on every call,
 * include/spin_ensembles/MonteCarloLoop.hpp:79-105): with on = 1, each call
 * after the first continues every chain from the configuration the previous
 * call left it in and skips the thermalisation sweeps, as long as the chain
 * count, N and the move type are unchanged.  RBM wavefunctions only. */
int rbmhip_mc_set_persistent(rbmhip_ens* ens, int on);
int rbmhip_mc_get_persistent(const rbmhip_ens* ens, int* on);
/* complex64 chains recompute theta = b + W^T s from the configuration every
 * `sweeps` sweeps (default 16; 0 never), bounding the rounding drift of the
 * incremental updates theta_j += 2 s'_p W_pj (include/quantum_state/Psi.hpp:113-119,
 * which the FP64 reference never re-synchronises).  complex128 chains follow
 * the reference and never re-synchronise. */
int rbmhip_mc_set_angle_resync(rbmhip_ens* ens, int sweeps);
/* Diagnostics of the incremental angle updates (no reference counterpart):
 * with on = 1 the RBM samplers keep each chain's final angles theta; read them
 * as chains x M complex (out may be NULL to query the sizes). */
int rbmhip_mc_record_angles(rbmhip_ens* ens, int on);
int rbmhip_mc_last_angles(const rbmhip_ens* ens, int64_t* chains, int* M, double* out);
/* num_steps (main.cpp:280, 290) */
int rbmhip_ens_num_steps(const rbmhip_ens* ens, uint64_t* out);
/* acceptance_rate counters of the last call (main.cpp:282-284) */
int rbmhip_mc_acceptance(const rbmhip_ens* ens, uint64_t* acceptances, uint64_t* rejections);
/* Per-sample records of the last Monte Carlo call (no reference counterpart:
 * the reference keeps no samples).  Any output may be NULL. */
int rbmhip_mc_last_samples(const rbmhip_ens* ens, uint64_t* spins, double* log_psi, double* e_loc);
int rbmhip_ens_destroy(rbmhip_ens* ens);

/* ---- batched evaluation -------------------------------------------------- */
/* log psi of n configurations (Psi::log_psi_s_std, source/quantum_state/Psi.cu:36-69) */
int rbmhip_log_psi(rbmhip_psi* psi, const uint64_t* spins, int64_t n, double* out);
/* local energy of n configurations (Operator::local_energy, include/operator/Operator.hpp:90-115) */
int rbmhip_local_energy(rbmhip_psi* psi, rbmhip_op* op, const uint64_t* spins, int64_t n, double* out);
/* Allocation-free local energies: the same computation and results as
 * rbmhip_local_energy, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The call allocates no device
 * memory: the model, the operator tables and the per-configuration records
 * live in a fixed 16 MiB device array per GPU that the library holds until the
 * process ends.  Limits: RBM wavefunctions (FP32 or FP64), n <= 262144, model
 * and operator tables <= 4 MiB; larger inputs fail with RBMHIP_ERR_ARG.  All
 * calls on one GPU share that array: calls that overlap on one GPU (from
 * several host threads) race and must be serialised by the caller. */
int rbmhip_local_energy_noalloc(rbmhip_psi* psi, rbmhip_op* op, const uint64_t* spins, int64_t n, double* out);
/* psi_O_k_vector(psi, spins): source/network_functions/PsiOkVector.cu:9-44;
 * writes num_params complex values, entries below 2N are zero */
int rbmhip_o_k_vector(rbmhip_psi* psi, const uint64_t* spins, double* out);
/* psi_vector: source/network_functions/PsiVector.cu:13-51 (2^N complex values) */
int rbmhip_psi_vector(rbmhip_psi* psi, double* out);
/* psi_norm: source/network_functions/PsiNorm.cu:13-44 (FP64 log-sum-exp) */
int rbmhip_psi_norm(rbmhip_psi* psi, rbmhip_ens* exact, double* out);

/* ---- network functions --------------------------------------------------- */
/* ExpectationValue::operator(): source/network_functions/ExpectationValue.cu:17-73 */
int rbmhip_expectation(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* out);
/* list version: ExpectationValue.cu:448-518; out has 2 * n_ops doubles */
int rbmhip_expectation_list(rbmhip_psi* psi, rbmhip_op* const* ops, int n_ops, rbmhip_ens* ens, double* out);
/* ExpectationValue::fluctuation: ExpectationValue.cu:130-175 */
int rbmhip_fluctuation(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* fluctuation, double* expectation);
/* ExpectationValue::gradient: ExpectationValue.cu:177-256; grad has num_params
 * complex values.  The reference's term conj<E_k> has the expectation
 * <E conj(O_k)> for a Hermitian operator; such operators use the centred
 * 2<E conj O> - 2<E> conj<O>, other operators the reference's E_k estimator. */
int rbmhip_gradient(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* grad, double* expectation);
/* The same gradient with a chosen estimator: RBMHIP_GRAD_EK is the reference's
 * <E conj O_k> + conj<E_k> - 2<E> conj<O_k> with E_k(s) = sum_n c_n psi(s'_n)/psi(s) O_k(s'_n)
 * (ExpectationValue.cu:177-256, Operator.hpp:145-179), valid for any operator;
 * RBMHIP_GRAD_HERMITIAN is 2<E conj O_k> - 2<E> conj<O_k> (Hermitian operators
 * only); RBMHIP_GRAD_AUTO (rbmhip_gradient) picks the Hermitian form when the
 * operator is Hermitian and E_k otherwise. */
#define RBMHIP_GRAD_AUTO 0
#define RBMHIP_GRAD_EK 1
#define RBMHIP_GRAD_HERMITIAN 2
int rbmhip_gradient_estimator(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, int estimator, double* grad,
                              double* expectation);
/* ExpectationValue::fluctuation_gradient: ExpectationValue.cu:258-332 (the
 * reference's E_k form); grad has num_params complex values */
int rbmhip_fluctuation_gradient(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* grad);
/* ExpectationValue::corrected: ExpectationValue.cu:74-127; value = <E>/<X>
 * (complex), norm = <X> with X = exp(2 Im E + Re(E^2 - E)).  As in the
 * reference, both local energies come from op; op2 is accepted and unused. */
int rbmhip_corrected(rbmhip_psi* psi, rbmhip_op* op, rbmhip_op* op2, rbmhip_ens* ens, double* value, double* norm);
/* ExpectationValue::difference: ExpectationValue.cu:334-446; on samples of psi,
 * out[i] = <E'_i> - <r E_i> / <r> with E'_i the local energy of psi_prime and
 * r = |psi/psi_prime|^2 (2 * n_ops doubles) */
int rbmhip_difference(rbmhip_psi* psi, rbmhip_psi* psi_prime, rbmhip_op* const* ops, int n_ops, rbmhip_ens* ens,
                      double* out);
/* get_S_matrix(Psi, ExactSummation): source/network_functions/S_matrix.cu:10-58
 * (num_params x num_params complex, unweighted and uncentred as in the reference) */
int rbmhip_s_matrix(rbmhip_psi* psi, rbmhip_ens* ens, double* out);
/* Centred, weighted S-matrix S_kl = <conj(O_k - <O_k>) (O_l - <O_l>)> over the
 * parameters that have an O_k (the num_params - 2N entries of `params` after
 * alpha and beta), P x P complex, FP32 MFMA products accumulated in FP64.  No
 * reference counterpart: the reference's S (get_S_matrix above) is unweighted
 * and uncentred and feeds no solver. */
int rbmhip_s_matrix_centred(rbmhip_psi* psi, rbmhip_ens* ens, double* out);
/* x = A^{-1} b for an n x n complex Hermitian positive-definite A (row-major,
 * lower triangle used), FP64 Cholesky on the GPU.  RBMHIP_ERR_NUMERIC if A is
 * not positive definite.  No reference counterpart. */
int rbmhip_cholesky_solve(int n, const double* A, const double* b, double* x);
/* Stochastic-reconfiguration step: update = (S + eps I)^{-1} g with S the
 * centred S-matrix and g the gradient of rbmhip_gradient, from one pass of the
 * ensemble; update has num_params complex values (zero for alpha and beta).
 * No reference counterpart (the reference's driver, pyRBMonGPU's
 * gradient_descent.py, has no S solve). */
int rbmhip_sr_update(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double eps, double* update, double* expectation);
/* HilbertSpaceDistance::distance / gradient (free_quantum_axis = false):
 * source/network_functions/HilbertSpaceDistance.cu:193-269.  grad (num_params
 * of psi_prime) is written when want_grad != 0. */
int rbmhip_hs_distance(rbmhip_psi* psi, rbmhip_psi* psi_prime, rbmhip_op* op, int is_unitary, rbmhip_ens* ens,
                       int want_grad, double nu, double* distance, double* grad);
/* The same with free_quantum_axis = true on psi (HilbertSpaceDistance.cu:47-80,
 * 109-122; update_quaxis 164-180): log psi' gains sum_i [(alpha'_i - alpha_i)
 * s_i r_i + i (beta'_i - beta_i)(r_i cos(alpha'_i + alpha_i) - s_i sin(...))]
 * with r_i = psi'(s with spin i flipped) / psi'(s), and grad entries i and N + i
 * are the alpha / beta derivatives.  alpha, beta are psi's axis angles,
 * alpha_prime, beta_prime psi_prime's (N each). */
int rbmhip_hs_distance_axis(rbmhip_psi* psi, rbmhip_psi* psi_prime, rbmhip_op* op, int is_unitary, rbmhip_ens* ens,
                            int want_grad, double nu, const double* alpha, const double* beta, const double* alpha_prime,
                            const double* beta_prime, double* distance, double* grad);

/* ---- learned proposal: masked autoregressive network (MADE) -------------- */
/* No counterpart in the reference.  A MADE (Germain et al. 2015) over x_i = 1
 * for spin up: a = W1 x + b1, h = relu(a), l = W2 h + b2,
 * q(x) = prod_j sigmoid(l_j)^x_j (1 - sigmoid(l_j))^(1 - x_j).  Hidden unit k has
 * degree m_k = 1 + k mod (N - 1); W1[k][i] is used iff i < m_k, W2[j][k] iff
 * m_k <= j, so l_j depends on x_0 .. x_{j-1}.  Parameters (FP32, flat):
 * [W1 (H x N), b1 (H), W2 (N x H), b2 (N)]; masked entries are kept at zero.
 * params == NULL draws the seeded initialisation: weight k (flat index) is
 * (2u - 1) / sqrt(fan_in), u = (x + 0.5) 2^-32 with x the first word of
 * Philox4x32-10(key = seed, ctr = (k, 0, 0, 3)); biases are zero.  N <= 128,
 * H <= 512, and 4 N round_up(H, 64) bytes (the bfloat16 weights) must fit in LDS. */
int rbmhip_made_create(int N, int H, const float* params, uint64_t seed, rbmhip_made** out);
int rbmhip_made_num_params(const rbmhip_made* made, int64_t* out);
int rbmhip_made_get_params(rbmhip_made* made, float* out);
int rbmhip_made_set_params(rbmhip_made* made, const float* params);
/* loss = -(1/n) sum_s log q(x_s) over the n configurations and its gradient
 * (num_params floats), FP32 MFMA forward and backward passes. */
int rbmhip_made_loss_grad(rbmhip_made* made, const uint64_t* spins, int64_t n, double* loss, float* grad);
/* `steps` Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) on minibatches of
 * `batch` configurations drawn with replacement from the n given ones (index
 * floor(u n), u from Philox(seed, ctr = (step, row, 0, 4))); batch == n uses all
 * n in order.  losses (may be NULL) receives the loss of each step's batch. */
int rbmhip_made_train(rbmhip_made* made, const uint64_t* spins, int64_t n, int steps, int batch, double lr,
                      uint64_t seed, double* losses);
int rbmhip_made_reset_optimizer(rbmhip_made* made);
/* Host-side check, without a GPU, that every kernel launch of one training
 * step (rbmhip_made_train / rbmhip_made_loss_grad) for this shape stays within
 * the buffer it is given; RBMHIP_ERR_ARG names the launch otherwise.  The
 * training calls run the same check before launching.  No reference
 * counterpart. */
int rbmhip_made_check_launches(int N, int H, int64_t n, int batch);
/* log q of n configurations with the sampler's arithmetic: bfloat16 weights,
 * FP32 hidden pre-activations, bfloat16 relu(a) in the logit dot products. */
int rbmhip_made_log_q(rbmhip_made* made, const uint64_t* spins, int64_t n, double* out);
/* n autoregressive samples and their log q (log_q may be NULL) */
int rbmhip_made_sample(rbmhip_made* made, int64_t n, uint64_t seed, uint64_t* spins, double* log_q);
int rbmhip_made_destroy(rbmhip_made* made);
/* Independent Metropolis-Hastings ensemble with the network as proposal; it is
 * used wherever a MonteCarloLoop is (expectation, gradient, distance).  Each
 * chain starts from a proposal, takes num_thermalization_steps unmeasured
 * proposals, then records one sample per proposal; a proposal s' is accepted
 * iff u <= |psi(s')|^2 q(s) / (|psi(s)|^2 q(s')).  The network must outlive
 * the ensemble. */
int rbmhip_learned_create(rbmhip_made* made, uint32_t num_samples, uint32_t num_thermalization_steps,
                          uint32_t num_markov_chains, uint64_t seed, rbmhip_ens** out);
/* Proposals of the last call: *n = chains * (num_thermalization_steps + steps + 1),
 * proposal t of chain c at index t * chains + c, with its log q and log psi
 * (complex).  Call with NULL arrays to read n first. */
int rbmhip_learned_last_proposals(const rbmhip_ens* ens, int64_t* n, uint64_t* spins, double* log_q, double* log_psi);
/* Number of measured configurations of the last call: a rejection repeats the
 * held configuration, and each run of equal samples is evaluated once and
 * weighted by its length. */
int rbmhip_learned_num_runs(const rbmhip_ens* ens, int64_t* out);

/* ---- kernel timing (HIP events on the library stream) -------------------- */
/* wait for all work on the library stream of the current device */
int rbmhip_synchronize(void);
int rbmhip_timing_enable(int on);
int rbmhip_timing_reset(void);
/* total milliseconds and launch count of a kernel family ("mc", "eval", "ok_reduce", ...) */
int rbmhip_timing_query(const char* name, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* RBMHIP_H */
