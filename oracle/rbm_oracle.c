/*
 * rbm_oracle.c -- FP64 CPU restatement of RBMonGPU's variational-Monte-Carlo
 * hot path.
 *
 * THIS FILE IS TEST INFRASTRUCTURE.  It is the checker for the HIP product
 * path in rbmongpu_amd/ and is only ever loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing in the
 * product links, loads or calls it.
 *
 * Every function restates the reference's own algorithm with its own
 * operation order wherever the order is observable in FP64, citing the file
 * and line it follows (paths relative to the reference repository root).
 * Limits are lifted: N <= 128 (two 64-bit spin words) and any M / layer width.
 *
 * Parity pins (each function is checked in tests/test_oracle_*.py against an
 * independent route: numpy closed forms, dense 2^N Hamiltonians built from
 * Kronecker products of Pauli matrices, central finite differences, Philox
 * known-answer vectors, exact-distribution statistics).  Functions with no
 * independent check are marked "parity unpinned" at their definition.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* Activation: include/quantum_state/psi_functions.hpp:49-57 (my_logcosh) and  */
/* 97-114 (my_tanh).  Branch on Re z > 0.0 exactly as the reference does.     */
/* ------------------------------------------------------------------------ */
static inline cplx lc(cplx z) {
    const double sign = creal(z) > 0.0 ? 1.0 : -1.0;
    return sign * 0.9003320053750442 * z + (
        5.49914721954 - sign * 2.16564366435 * z
    ) / (
        9.19376335670885 + z * (sign * 10.2180213465 + z * (7.771429504240965 + z * (sign * 3.746646023906276 + z)))
    ) - 0.598139;
}

static inline cplx th(cplx z) {
    const double sign = creal(z) > 0.0 ? 1.0 : -1.0;
    const cplx denominator = 9.19376335670885 + z * (sign * 10.218021346543315 + z * (7.771429504240965 + z * (sign * 3.746646023906276 + z)));
    return (
        z * (
            83.68563506532087 + z * (
                sign * 177.6769746361748 + z * (
                    199.24474920889975 + z * (
                        sign * 146.36284300074402 + z * (
                            70.82878897882324 + z * (
                                sign * 26.632014683761202 + z * (
                                    6.746450656267947 + sign * 0.9003320053750442 * z
        )))))))
    ) / (denominator * denominator);
}

void orc_logcosh(const cplx* z, cplx* out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = lc(z[i]); }
void orc_tanh(const cplx* z, cplx* out, int64_t n) { for (int64_t i = 0; i < n; i++) out[i] = th(z[i]); }

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The reference draws   */
/* from curand XORWOW (include/random.h:21-54); that stream cannot be        */
/* reproduced without CUDA, so the rebuilt sampler and this oracle share one */
/* counter-based generator instead.  Pinned by the Random123 known-answer    */
/* vectors in tests/test_oracle_rng.py.                                      */
/* ------------------------------------------------------------------------ */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; r++) {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    memcpy(out, c, sizeof(c));
}

/* ------------------------------------------------------------------------ */
/* Spins: include/Spins.h:125-129 -- spin i = +1 iff bit i is set.            */
/* ------------------------------------------------------------------------ */
static inline double spin(const uint64_t* s, int i) {
    return ((s[i >> 6] >> (i & 63)) & 1u) ? 1.0 : -1.0;
}
static inline void flip(uint64_t* s, int i) { s[i >> 6] ^= (uint64_t)1 << (i & 63); }
static inline void mask_n(uint64_t* s, int N) {
    /* Spins(config, N) masks to N bits: include/Spins.h:230-237 */
    if (N < 64) { s[0] &= (((uint64_t)1) << N) - 1u; s[1] = 0; }
    else if (N == 64) { s[1] = 0; }
    else if (N < 128) { s[1] &= (((uint64_t)1) << (N - 64)) - 1u; }
}

/* rotate_left(shift, N) for N <= 64 (include/Spins.h:280-285); generic two-  */
/* word version for 64 < N <= 128 follows the same definition bit by bit.     */
static void rotate_left(const uint64_t* s, int shift, int N, uint64_t* out) {
    if (N <= 64) {
        uint64_t c = s[0];
        uint64_t r = shift == 0 ? c : ((c << shift) | (c >> (N - shift)));
        out[0] = r; out[1] = 0; mask_n(out, N);
        return;
    }
    out[0] = out[1] = 0;
    for (int i = 0; i < N; i++) {
        if ((s[i >> 6] >> (i & 63)) & 1u) {
            int j = (i + shift) % N;
            out[j >> 6] |= (uint64_t)1 << (j & 63);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Model descriptors                                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t kind;            /* 0 = RBM (Psi), 1 = PsiDeep */
    int32_t N;
    double  prefactor;
    /* RBM: include/quantum_state/Psi.hpp:31-53 */
    int32_t M;
    const cplx* b;           /* M */
    const cplx* W;           /* N x M row-major, W[i*M + j] */
    const cplx* a;           /* visible bias extension (N) or NULL */
    /* PsiDeep: include/quantum_state/PsiDeep.hpp:39-84, flattened */
    int32_t num_layers;
    const int32_t* sizes;        /* num_layers */
    const int32_t* conn_counts;  /* lhs_connectivity per layer */
    const uint32_t* conns;       /* concatenated C_l x size_l */
    const cplx* weights;         /* concatenated C_l x size_l */
    const cplx* biases;          /* concatenated size_l */
    const cplx* final_weights;   /* size_{L-1} */
    int32_t translational_invariance;
    int32_t ref_compat_final;    /* 1: final-weight O_k = log_psi * x_j (PsiDeep.hpp:285-291) */
} orc_model;

typedef struct {
    int32_t S, L;
    const cplx* coeff;       /* S */
    const int32_t* types;    /* S x L, 0=I 1=X 2=Y 3=Z (include/operator/MatrixElement.hpp:10-12) */
    const int32_t* idx;      /* S x L, -1 terminated */
} orc_op;

static int64_t rbm_num_params(const orc_model* m) {
    return 2 * (int64_t)m->N + m->M + (int64_t)m->N * m->M + (m->a ? m->N : 0);
}

static int64_t deep_num_params(const orc_model* m) {
    /* source/quantum_state/PsiDeep.cu:36-61 */
    int64_t p = 2 * (int64_t)m->N;
    for (int l = 0; l < m->num_layers; l++) p += m->sizes[l] + (int64_t)m->conn_counts[l] * m->sizes[l];
    p += m->sizes[m->num_layers - 1];
    return p;
}

int64_t orc_num_params(const orc_model* m) { return m->kind == 0 ? rbm_num_params(m) : deep_num_params(m); }

/* ---------------------------- RBM ------------------------------------- */
/* Psi::angle: include/quantum_state/Psi.hpp:43-53 */
static void rbm_angles(const orc_model* m, const uint64_t* s, cplx* theta) {
    for (int j = 0; j < m->M; j++) {
        cplx r = m->b[j];
        for (int i = 0; i < m->N; i++) r += m->W[(int64_t)i * m->M + j] * spin(s, i);
        theta[j] = r;
    }
}

/* Visible-bias extension: sum_i a_i s_i (no counterpart in the reference). */
static cplx rbm_visible(const orc_model* m, const uint64_t* s) {
    cplx r = 0;
    if (!m->a) return 0;
    for (int i = 0; i < m->N; i++) r += m->a[i] * spin(s, i);
    return r;
}

/* Psi::log_psi_s (CPU branch): include/quantum_state/Psi.hpp:80-86 */
static cplx rbm_log_psi_from_angles(const orc_model* m, const uint64_t* s, const cplx* theta) {
    cplx r = 0;
    for (int j = 0; j < m->M; j++) r += lc(theta[j]);
    return r + rbm_visible(m, s);
}

/* Psi::flip_spin_of_jth_angle: include/quantum_state/Psi.hpp:113-119 */
static void rbm_flip_angles(const orc_model* m, int position, const uint64_t* new_spins, cplx* theta) {
    const double sp = spin(new_spins, position);
    for (int j = 0; j < m->M; j++) theta[j] += 2.0 * sp * m->W[(int64_t)position * m->M + j];
}

/* Psi::foreach_O_k / get_O_k_element: include/quantum_state/Psi.hpp:165-202 */
static void rbm_o_k(const orc_model* m, const uint64_t* s, const cplx* theta, cplx* out) {
    const int N = m->N, M = m->M;
    for (int k = 0; k < 2 * N; k++) out[k] = 0;
    for (int j = 0; j < M; j++) out[2 * N + j] = th(theta[j]);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < M; j++) out[2 * N + M + (int64_t)i * M + j] = out[2 * N + j] * spin(s, i);
    if (m->a) {
        const int64_t base = 2 * (int64_t)N + M + (int64_t)N * M;
        for (int i = 0; i < N; i++) out[base + i] = spin(s, i);
    }
}

/* ---------------------------- PsiDeep --------------------------------- */
static int64_t deep_layer_offset(const orc_model* m, int layer, int which) {
    /* which 0: conns/weights offset, 1: biases offset */
    int64_t off = 0;
    for (int l = 0; l < layer; l++) off += which == 0 ? (int64_t)m->conn_counts[l] * m->sizes[l] : m->sizes[l];
    return off;
}

static int deep_max_width(const orc_model* m) {
    int w = m->N;
    for (int l = 0; l < m->num_layers; l++) if (m->sizes[l] > w) w = m->sizes[l];
    return w;
}

static int deep_num_units(const orc_model* m) {
    int u = 0;
    for (int l = 0; l < m->num_layers; l++) u += m->sizes[l];
    return u;
}

/* PsiDeepT::forward_pass: include/quantum_state/PsiDeep.hpp:89-133.
 * Stores pre-activations in deep_angles (if non-NULL) and returns
 * sum_j x_j * final_weights_j of the last layer. */
static cplx deep_forward(const orc_model* m, const uint64_t* s, cplx* act, cplx* deep_angles) {
    const int W = deep_max_width(m);
    cplx* reg = (cplx*)malloc(sizeof(cplx) * (size_t)W);
    for (int i = 0; i < m->N; i++) act[i] = spin(s, i);
    int begin_angles = 0;
    for (int l = 0; l < m->num_layers; l++) {
        const int size = m->sizes[l], C = m->conn_counts[l];
        const uint32_t* conn = m->conns + deep_layer_offset(m, l, 0);
        const cplx* w = m->weights + deep_layer_offset(m, l, 0);
        const cplx* b = m->biases + deep_layer_offset(m, l, 1);
        for (int j = 0; j < size; j++) {
            cplx r = 0;
            for (int i = 0; i < C; i++) r += w[(int64_t)i * size + j] * act[conn[(int64_t)i * size + j]];
            r += b[j];
            reg[j] = r;
            if (deep_angles) deep_angles[begin_angles + j] = r;
        }
        for (int k = 0; k < size; k++) act[k] = lc(reg[k]);
        begin_angles += size;
    }
    cplx result = 0;
    const int last = m->sizes[m->num_layers - 1];
    for (int j = 0; j < last; j++) result += act[j] * m->final_weights[j];
    free(reg);
    return result;
}

/* PsiDeepT::log_psi_s_generic: include/quantum_state/PsiDeep.hpp:135-169 */
static cplx deep_log_psi(const orc_model* m, const uint64_t* s) {
    cplx* act = (cplx*)malloc(sizeof(cplx) * (size_t)deep_max_width(m));
    cplx result = 0;
    const int shifts = m->translational_invariance ? m->N : 1;
    for (int shift = 0; shift < shifts; shift++) {
        uint64_t sh[2];
        rotate_left(s, shift, m->N, sh);
        result += deep_forward(m, sh, act, NULL);
    }
    if (m->translational_invariance) result *= 1.0 / m->N;
    free(act);
    return result;
}

/* PsiDeepT::foreach_O_k: include/quantum_state/PsiDeep.hpp:214-293.  Uses the
 * un-shifted spins (the reference ignores translational invariance here).
 * The rhs connectivity (source/quantum_state/PsiDeep.cu:86-117) is the
 * transpose of the lhs connections; it is applied here by scattering over the
 * lhs table, which visits the same (unit, weight) pairs. */
static void deep_o_k(const orc_model* m, const uint64_t* s, cplx* out) {
    const int L = m->num_layers, Wd = deep_max_width(m);
    const int64_t P = deep_num_params(m);
    cplx* act = (cplx*)malloc(sizeof(cplx) * (size_t)Wd);
    cplx* delta = (cplx*)malloc(sizeof(cplx) * (size_t)Wd);
    cplx* nd = (cplx*)malloc(sizeof(cplx) * (size_t)Wd);
    cplx* ang = (cplx*)malloc(sizeof(cplx) * (size_t)deep_num_units(m));
    for (int64_t k = 0; k < P; k++) out[k] = 0;
    const cplx log_psi = deep_forward(m, s, act, ang);

    int64_t* begin_params = (int64_t*)malloc(sizeof(int64_t) * (size_t)L);
    int* begin_angles = (int*)malloc(sizeof(int) * (size_t)L);
    int64_t p = 2 * (int64_t)m->N; int a = 0;
    for (int l = 0; l < L; l++) {
        begin_params[l] = p; begin_angles[l] = a;
        p += m->sizes[l] + (int64_t)m->conn_counts[l] * m->sizes[l];
        a += m->sizes[l];
    }
    for (int l = L - 1; l >= 0; l--) {
        const int size = m->sizes[l], C = m->conn_counts[l];
        const uint32_t* conn = m->conns + deep_layer_offset(m, l, 0);
        if (l == L - 1) {
            for (int j = 0; j < size; j++) delta[j] = m->final_weights[j] * th(ang[begin_angles[l] + j]);
        } else {
            /* delta_i = tanh(a_i) * sum_{j in next layer, conn(., j) = i} W_next[., j] delta_next[j] */
            const int nsize = m->sizes[l + 1], nC = m->conn_counts[l + 1];
            const uint32_t* nconn = m->conns + deep_layer_offset(m, l + 1, 0);
            const cplx* nw = m->weights + deep_layer_offset(m, l + 1, 0);
            for (int i = 0; i < size; i++) nd[i] = 0;
            for (int j = 0; j < nsize; j++)
                for (int i = 0; i < nC; i++) {
                    const uint32_t u = nconn[(int64_t)i * nsize + j];
                    nd[u] += nw[(int64_t)i * nsize + j] * delta[j];
                }
            for (int i = 0; i < size; i++) nd[i] *= th(ang[begin_angles[l] + i]);
            for (int i = 0; i < size; i++) delta[i] = nd[i];
        }
        for (int j = 0; j < size; j++) {
            out[begin_params[l] + j] = delta[j];
            for (int i = 0; i < C; i++) {
                const uint32_t u = conn[(int64_t)i * size + j];
                const cplx x = l == 0 ? (cplx)spin(s, (int)u) : lc(ang[begin_angles[l - 1] + u]);
                out[begin_params[l] + size + (int64_t)i * size + j] = delta[j] * x;
            }
        }
    }
    const int last = m->sizes[L - 1];
    for (int j = 0; j < last; j++) {
        const cplx x = lc(ang[begin_angles[L - 1] + j]);
        out[P - last + j] = m->ref_compat_final ? log_psi * x : x;
    }
    free(act); free(delta); free(nd); free(ang); free(begin_params); free(begin_angles);
}

/* ---------------------------- generic model --------------------------- */
typedef struct {
    const orc_model* m;
    cplx* theta;     /* RBM angle cache (M) */
} state_t;

static cplx model_log_psi(const orc_model* m, const uint64_t* s) {
    if (m->kind == 0) {
        cplx* th_ = (cplx*)malloc(sizeof(cplx) * (size_t)m->M);
        rbm_angles(m, s, th_);
        cplx r = rbm_log_psi_from_angles(m, s, th_);
        free(th_);
        return r;
    }
    return deep_log_psi(m, s);
}

void orc_log_psi(const orc_model* m, const uint64_t* spins, int64_t n, cplx* out) {
    for (int64_t k = 0; k < n; k++) out[k] = model_log_psi(m, spins + 2 * k);
}

/* psi_O_k_vector (source/network_functions/PsiOkVector.cu:9-44) writes O_k at
 * indices 2N+k; entries below 2N are zero here (uninitialised in the
 * reference). out has num_params entries. */
void orc_o_k(const orc_model* m, const uint64_t* s, cplx* out) {
    if (m->kind == 0) {
        cplx* th_ = (cplx*)malloc(sizeof(cplx) * (size_t)m->M);
        rbm_angles(m, s, th_);
        rbm_o_k(m, s, th_, out);
        free(th_);
    } else {
        deep_o_k(m, s, out);
    }
}

/* ---------------------------- Operator -------------------------------- */
/* Operator::nth_matrix_element: include/operator/Operator.hpp:45-88.  For the
 * RBM, angles are flipped incrementally (flip_spin_of_jth_angle); PsiDeep has
 * no angle cache (PsiDeep.hpp:191-194). Y and Z factors use the ORIGINAL spins. */
static cplx nth_matrix_element(const orc_model* m, const orc_op* op, const uint64_t* s, int n,
                               uint64_t* s_out, cplx* theta /* may be NULL */) {
    cplx coef = op->coeff[n];
    s_out[0] = s[0]; s_out[1] = s[1];
    for (int64_t t = (int64_t)n * op->L; ; t++) {
        const int pidx = op->idx[t];
        if (pidx == -1) break;
        const int pt = op->types[t];
        if (pt == 1) {
            flip(s_out, pidx);
            if (theta && m->kind == 0) rbm_flip_angles(m, pidx, s_out, theta);
        } else if (pt == 2) {
            flip(s_out, pidx);
            if (theta && m->kind == 0) rbm_flip_angles(m, pidx, s_out, theta);
            coef *= (0.0 - 1.0 * I) * spin(s, pidx);
        } else if (pt == 3) {
            coef *= spin(s, pidx);
        }
    }
    return coef;
}

static cplx log_psi_prime(const orc_model* m, const uint64_t* sp, const cplx* theta_p) {
    if (m->kind == 0) return rbm_log_psi_from_angles(m, sp, theta_p);
    return deep_log_psi(m, sp);
}

/* Operator::local_energy: include/operator/Operator.hpp:90-115 */
static cplx local_energy(const orc_model* m, const orc_op* op, const uint64_t* s, cplx log_psi, const cplx* theta) {
    cplx result = 0;
    cplx* tp = m->kind == 0 ? (cplx*)malloc(sizeof(cplx) * (size_t)m->M) : NULL;
    for (int n = 0; n < op->S; n++) {
        uint64_t sp[2];
        if (tp) memcpy(tp, theta, sizeof(cplx) * (size_t)m->M);
        const cplx coef = nth_matrix_element(m, op, s, n, sp, tp);
        const cplx lpp = log_psi_prime(m, sp, tp);
        result += coef * cexp(lpp - log_psi);
    }
    free(tp);
    return result;
}

/* Operator::foreach_E_k_s_prime: include/operator/Operator.hpp:145-179;
 * accumulates weight * E_s' * O_k(s') into acc (num_params). */
static void accumulate_E_k2(const orc_model* m, const orc_op* op, const uint64_t* s, cplx log_psi,
                            const cplx* theta, double weight, cplx* acc, cplx weight2, cplx* acc2, cplx* scratch_ok);
static void accumulate_E_k(const orc_model* m, const orc_op* op, const uint64_t* s, cplx log_psi,
                           const cplx* theta, double weight, cplx* acc, cplx* scratch_ok) {
    accumulate_E_k2(m, op, s, log_psi, theta, weight, acc, 0, NULL, scratch_ok);
}

/* acc += weight E_k(s), acc2 += weight2 E_k(s) (acc2 may be NULL):
 * Operator.hpp:145-179 */
static void accumulate_E_k2(const orc_model* m, const orc_op* op, const uint64_t* s, cplx log_psi,
                            const cplx* theta, double weight, cplx* acc, cplx weight2, cplx* acc2, cplx* scratch_ok) {
    const int64_t P = orc_num_params(m);
    cplx* tp = m->kind == 0 ? (cplx*)malloc(sizeof(cplx) * (size_t)m->M) : NULL;
    for (int n = 0; n < op->S; n++) {
        uint64_t sp[2];
        if (tp) memcpy(tp, theta, sizeof(cplx) * (size_t)m->M);
        const cplx coef = nth_matrix_element(m, op, s, n, sp, tp);
        const cplx lpp = log_psi_prime(m, sp, tp);
        const cplx e = coef * cexp(lpp - log_psi);
        if (m->kind == 0) rbm_o_k(m, sp, tp, scratch_ok); else deep_o_k(m, sp, scratch_ok);
        for (int64_t k = 2 * (int64_t)m->N; k < P; k++) acc[k] += weight * (e * scratch_ok[k]);
        if (acc2)
            for (int64_t k = 2 * (int64_t)m->N; k < P; k++) acc2[k] += weight2 * (e * scratch_ok[k]);
    }
    free(tp);
}

void orc_local_energy(const orc_model* m, const orc_op* op, const uint64_t* spins, int64_t n, cplx* out) {
    cplx* theta = m->kind == 0 ? (cplx*)malloc(sizeof(cplx) * (size_t)m->M) : NULL;
    for (int64_t k = 0; k < n; k++) {
        const uint64_t* s = spins + 2 * k;
        if (theta) rbm_angles(m, s, theta);
        const cplx lp = m->kind == 0 ? rbm_log_psi_from_angles(m, s, theta) : deep_log_psi(m, s);
        out[k] = local_energy(m, op, s, lp, theta);
    }
    free(theta);
}

/* ------------------------------------------------------------------------ */
/* Spin ensembles                                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t kind;              /* 0 = ExactSummation, 1 = MonteCarloLoop */
    int32_t N;
    /* ExactSummation: include/spin_ensembles/ExactSummation.hpp:36-84 */
    int64_t num_configs;       /* 2^N, or the sector size */
    const uint64_t* configs;   /* sector list (2 words each) or NULL */
    /* MonteCarloLoop: include/spin_ensembles/MonteCarloLoop.hpp:55-256 */
    uint32_t num_samples, num_sweeps, num_thermalization_sweeps, num_markov_chains;
    uint64_t seed;
    uint32_t call_index;
    /* outputs */
    uint64_t acceptances, rejections;
    /* MonteCarloLoop move (MonteCarloLoop.hpp:79-93, 153-201): 0 single flip,
     * 1 exchange conserving total S_z with k_up spins up, 2 fast sweep with
     * `tries` random masks ANDed */
    int32_t move, tries, k_up;
} orc_ens;

typedef void (*sample_fn)(void* ctx, int64_t index, const uint64_t* s, cplx log_psi, const cplx* theta, double weight);

static int64_t ens_num_steps(const orc_ens* e) {
    return e->kind == 0 ? e->num_configs : (int64_t)e->num_samples;
}

int64_t orc_ens_num_steps(const orc_ens* e) { return ens_num_steps(e); }

/* Probability: include/quantum_state/Psi.hpp:133-136 */
static double probability(const orc_model* m, double log_psi_real) {
    return exp(2.0 * (log(m->prefactor) + log_psi_real));
}

/* ExactSummation::kernel_foreach (CPU branch): ExactSummation.hpp:47-84 */
static void exact_foreach(const orc_model* m, orc_ens* e, sample_fn f, void* ctx) {
    cplx* theta = m->kind == 0 ? (cplx*)malloc(sizeof(cplx) * (size_t)m->M) : NULL;
    for (int64_t idx = 0; idx < e->num_configs; idx++) {
        uint64_t s[2];
        if (e->configs) { s[0] = e->configs[2 * idx]; s[1] = e->configs[2 * idx + 1]; }
        else { s[0] = (uint64_t)idx; s[1] = 0; mask_n(s, e->N); }
        if (theta) rbm_angles(m, s, theta);
        const cplx lp = m->kind == 0 ? rbm_log_psi_from_angles(m, s, theta) : deep_log_psi(m, s);
        const double weight = (double)e->num_configs * probability(m, creal(lp));
        f(ctx, idx, s, lp, theta, weight);
    }
    free(theta);
}

/* Random numbers of one proposal: proposal k of chain c in call `call` uses
 * Philox(key = seed, ctr = (c, k >> 1, call, 0)); words (0,1) serve even k,
 * words (2,3) odd k.  Word x picks the position x % N (MonteCarloLoop.hpp:175)
 * and word y gives u = (y + 0.5) * 2^-32 in (0, 1) for the acceptance test. */
static void proposal_randoms(const orc_ens* e, uint32_t chain, uint64_t k, uint32_t* x, double* u) {
    const uint32_t key[2] = {(uint32_t)e->seed, (uint32_t)(e->seed >> 32)};
    const uint32_t ctr[4] = {chain, (uint32_t)(k >> 1), e->call_index, 0u};
    uint32_t r[4];
    orc_philox(ctr, key, r);
    const uint32_t xx = (k & 1) ? r[2] : r[0];
    const uint32_t yy = (k & 1) ? r[3] : r[1];
    *x = xx;
    *u = ((double)yy + 0.5) * (1.0 / 4294967296.0);
}

/* Initial configuration of a chain (Spins::random, include/Spins.h:87-104):
 * 128 random bits from Philox(key = seed, ctr = (c, 0, call, 1)), masked to N. */
static void initial_spins(const orc_ens* e, uint32_t chain, uint64_t* s) {
    const uint32_t key[2] = {(uint32_t)e->seed, (uint32_t)(e->seed >> 32)};
    const uint32_t ctr[4] = {chain, 0u, e->call_index, 1u};
    uint32_t r[4];
    orc_philox(ctr, key, r);
    s[0] = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
    s[1] = (uint64_t)r[2] | ((uint64_t)r[3] << 32);
    mask_n(s, e->N);
}

/* Move draws, identical to the library's propose_move / sector_start:
 * exchange: Philox(c, k, call, 4) = (x, y, z): flip p = x % N, then flip the
 * (y % cnt)-th (in index order) of the cnt spins equal to the new value of p
 * (p included); u from z.  Fast sweep: mask = AND of max(tries, 1) N-bit
 * strings Philox(c, k, call, 0x10 + j); u from Philox(c, k, call, 5).x.
 * Sector start: flip single bits chosen by Philox(c, j, call, 6).x among the
 * set (too many up) or clear (too few) bits until k_up spins are up. */
static int select_bit_c(const uint64_t* w, int k) {
    int base = 0;
    uint64_t x = w[0];
    const int c0 = __builtin_popcountll(w[0]);
    if (k >= c0) { k -= c0; x = w[1]; base = 64; }
    for (int i = 0; i < k; i++) x &= x - 1u;
    return base + __builtin_ctzll(x);
}

static void full_mask_c(int N, uint64_t* m) {
    m[0] = ~(uint64_t)0; m[1] = ~(uint64_t)0;
    mask_n(m, N);
}

static void sector_start_c(const orc_ens* e, uint32_t chain, int N, uint64_t* s) {
    const uint32_t key[2] = {(uint32_t)e->seed, (uint32_t)(e->seed >> 32)};
    uint64_t m[2];
    full_mask_c(N, m);
    for (uint32_t j = 0;; j++) {
        const int up = __builtin_popcountll(s[0]) + __builtin_popcountll(s[1]);
        if (up == e->k_up) break;
        const uint32_t ctr[4] = {chain, j, e->call_index, 6u};
        uint32_t r[4];
        orc_philox(ctr, key, r);
        if (up > e->k_up) {
            flip(s, select_bit_c(s, (int)(r[0] % (uint32_t)up)));
        } else {
            const uint64_t c[2] = {~s[0] & m[0], ~s[1] & m[1]};
            flip(s, select_bit_c(c, (int)(r[0] % (uint32_t)(N - up))));
        }
    }
}

static void propose_move_c(const orc_ens* e, int N, uint32_t chain, uint64_t k, const uint64_t* s, uint64_t* n, double* u) {
    const uint32_t key[2] = {(uint32_t)e->seed, (uint32_t)(e->seed >> 32)};
    uint64_t m[2];
    full_mask_c(N, m);
    uint32_t r[4];
    if (e->move == 1) {
        const uint32_t ctr[4] = {chain, (uint32_t)k, e->call_index, 4u};
        orc_philox(ctr, key, r);
        n[0] = s[0]; n[1] = s[1];
        const int p = (int)(r[0] % (uint32_t)N);
        flip(n, p);
        const int v = (int)((p < 64 ? n[0] >> p : n[1] >> (p - 64)) & 1u);
        const uint64_t c[2] = {v ? n[0] : (~n[0] & m[0]), v ? n[1] : (~n[1] & m[1])};
        const int cnt = __builtin_popcountll(c[0]) + __builtin_popcountll(c[1]);
        flip(n, select_bit_c(c, (int)(r[1] % (uint32_t)cnt)));
        *u = ((double)r[2] + 0.5) * (1.0 / 4294967296.0);
    } else {
        uint64_t f[2] = {m[0], m[1]};
        const int nt = e->tries > 1 ? e->tries : 1;
        for (int j = 0; j < nt; j++) {
            const uint32_t ctr[4] = {chain, (uint32_t)k, e->call_index, 0x10u + (uint32_t)j};
            orc_philox(ctr, key, r);
            f[0] &= (uint64_t)r[0] | ((uint64_t)r[1] << 32);
            f[1] &= (uint64_t)r[2] | ((uint64_t)r[3] << 32);
        }
        n[0] = s[0] ^ f[0]; n[1] = s[1] ^ f[1];
        const uint32_t ctr[4] = {chain, (uint32_t)k, e->call_index, 5u};
        orc_philox(ctr, key, r);
        *u = ((double)r[0] + 0.5) * (1.0 / 4294967296.0);
    }
}

static double model_log_psi_real_state(const orc_model* m, const uint64_t* s, const cplx* theta) {
    if (m->kind == 0) return creal(rbm_log_psi_from_angles(m, s, theta));
    return creal(deep_log_psi(m, s));
}

/* MonteCarloLoop::mc_update: include/spin_ensembles/MonteCarloLoop.hpp:153-256
 * (single flip; the exchange and fast-sweep moves draw as propose_move_c). */
static void mc_update(const orc_model* m, orc_ens* e, uint32_t chain, uint64_t k,
                      double* log_psi_real, uint64_t* s, cplx* theta) {
    if (e->move != 0) {
        /* exchange / fast sweep: the angles of the proposal from scratch */
        uint64_t n[2];
        double u;
        propose_move_c(e, m->N, chain, k, s, n, &u);
        cplx* tn = NULL;
        if (m->kind == 0) {
            tn = (cplx*)malloc(sizeof(cplx) * (size_t)m->M);
            rbm_angles(m, n, tn);
        }
        const double next = model_log_psi_real_state(m, n, tn);
        const double ratio = exp(2.0 * (next - *log_psi_real));
        if (ratio > 1.0 || u <= ratio) {
            *log_psi_real = next;
            s[0] = n[0]; s[1] = n[1];
            if (tn) memcpy(theta, tn, sizeof(cplx) * (size_t)m->M);
            e->acceptances++;
        } else {
            e->rejections++;
        }
        free(tn);
        return;
    }
    uint32_t x; double u;
    proposal_randoms(e, chain, k, &x, &u);
    const int position = (int)(x % (uint32_t)m->N);
    flip(s, position);
    if (m->kind == 0) rbm_flip_angles(m, position, s, theta);
    const double next = model_log_psi_real_state(m, s, theta);
    const double ratio = exp(2.0 * (next - *log_psi_real));
    if (ratio > 1.0 || u <= ratio) {
        *log_psi_real = next;
        e->acceptances++;
    } else {
        e->rejections++;
        flip(s, position);
        if (m->kind == 0) rbm_flip_angles(m, position, s, theta);
    }
}

/* MonteCarloLoop::kernel_foreach: MonteCarloLoop.hpp:55-138.  Chains are run
 * one after another; sample index = step * num_markov_chains + chain. */
static void mc_foreach(const orc_model* m, orc_ens* e, sample_fn f, void* ctx) {
    const uint32_t chains = e->num_markov_chains;
    const uint32_t steps = e->num_samples / chains;
    cplx* theta = m->kind == 0 ? (cplx*)malloc(sizeof(cplx) * (size_t)m->M) : NULL;
    for (uint32_t c = 0; c < chains; c++) {
        uint64_t s[2];
        initial_spins(e, c, s);
        if (e->move == 1) sector_start_c(e, c, m->N, s);
        if (theta) rbm_angles(m, s, theta);
        double lpr = model_log_psi_real_state(m, s, theta);
        uint64_t k = 0;
        /* one sample per num_sweeps proposals with the fast sweep (MonteCarloLoop.hpp:109-112) */
        const uint64_t per_sweep = e->move == 2 ? 1u : (uint64_t)m->N;
        for (uint64_t i = 0; i < (uint64_t)e->num_thermalization_sweeps * m->N; i++) mc_update(m, e, c, k++, &lpr, s, theta);
        for (uint32_t step = 0; step < steps; step++) {
            for (uint64_t i = 0; i < (uint64_t)e->num_sweeps * per_sweep; i++) mc_update(m, e, c, k++, &lpr, s, theta);
            const cplx lp = m->kind == 0 ? rbm_log_psi_from_angles(m, s, theta) : deep_log_psi(m, s);
            f(ctx, (int64_t)step * chains + c, s, lp, theta, 1.0);
        }
    }
    free(theta);
}

static void ens_foreach(const orc_model* m, orc_ens* e, sample_fn f, void* ctx) {
    if (e->kind == 0) exact_foreach(m, e, f, ctx); else mc_foreach(m, e, f, ctx);
}

/* ------------------------------------------------------------------------ */
/* Network functions                                                        */
/* ------------------------------------------------------------------------ */

/* Per-sample records: spins, log psi, E_loc (op may be NULL), weight. */
typedef struct {
    const orc_model* m; const orc_op* op;
    uint64_t* spins; cplx* log_psi; cplx* e_loc; double* weight;
} rec_ctx;

static void rec_fn(void* vctx, int64_t idx, const uint64_t* s, cplx lp, const cplx* theta, double w) {
    rec_ctx* c = (rec_ctx*)vctx;
    if (c->spins) { c->spins[2 * idx] = s[0]; c->spins[2 * idx + 1] = s[1]; }
    if (c->log_psi) c->log_psi[idx] = lp;
    if (c->e_loc) c->e_loc[idx] = c->op ? local_energy(c->m, c->op, s, lp, theta) : 0;
    if (c->weight) c->weight[idx] = w;
}

void orc_records(const orc_model* m, const orc_op* op, orc_ens* e,
                 uint64_t* spins, cplx* log_psi, cplx* e_loc, double* weight) {
    rec_ctx c = {m, op, spins, log_psi, e_loc, weight};
    ens_foreach(m, e, rec_fn, &c);
}

/* ExpectationValue::operator(): source/network_functions/ExpectationValue.cu:17-73 */
typedef struct { const orc_model* m; const orc_op* op; cplx acc; double acc2; } ev_ctx;
static void ev_fn(void* vctx, int64_t idx, const uint64_t* s, cplx lp, const cplx* theta, double w) {
    ev_ctx* c = (ev_ctx*)vctx; (void)idx;
    const cplx el = local_energy(c->m, c->op, s, lp, theta);
    c->acc += w * el;
    c->acc2 += w * (creal(el) * creal(el) + cimag(el) * cimag(el));
}

void orc_expectation(const orc_model* m, const orc_op* op, orc_ens* e, cplx* out) {
    ev_ctx c = {m, op, 0, 0};
    ens_foreach(m, e, ev_fn, &c);
    *out = c.acc * (1.0 / ens_num_steps(e));
}

/* ExpectationValue::fluctuation: ExpectationValue.cu:130-175 */
void orc_fluctuation(const orc_model* m, const orc_op* op, orc_ens* e, double* fluct, cplx* E) {
    ev_ctx c = {m, op, 0, 0};
    ens_foreach(m, e, ev_fn, &c);
    const cplx El = c.acc * (1.0 / ens_num_steps(e));
    const double El2 = c.acc2 * (1.0 / ens_num_steps(e));
    *E = El;
    *fluct = sqrt(El2 - (creal(El) * creal(El) + cimag(El) * cimag(El)));
}

/* ExpectationValue::gradient: ExpectationValue.cu:177-256.
 * mode 0: the reference estimator  <E O*> + conj<E_k> - 2 <E> conj<O>;
 * mode 1: 2 <E O*> - 2 <E> conj<O> (equal to mode 0 under ExactSummation for
 * a Hermitian operator; pinned against mode 0 in the tests). */
typedef struct {
    const orc_model* m; const orc_op* op; int mode; int64_t P;
    cplx E; cplx* O; cplx* EO; cplx* Ek; cplx* scratch;
} gr_ctx;
static void gr_fn(void* vctx, int64_t idx, const uint64_t* s, cplx lp, const cplx* theta, double w) {
    gr_ctx* c = (gr_ctx*)vctx; (void)idx;
    const cplx el = local_energy(c->m, c->op, s, lp, theta);
    c->E += w * el;
    if (c->m->kind == 0) rbm_o_k(c->m, s, theta, c->scratch); else deep_o_k(c->m, s, c->scratch);
    for (int64_t k = 2 * (int64_t)c->m->N; k < c->P; k++) {
        c->O[k] += w * c->scratch[k];
        c->EO[k] += w * el * conj(c->scratch[k]);
    }
    if (c->mode == 0) accumulate_E_k(c->m, c->op, s, lp, theta, w, c->Ek, c->scratch);
}

void orc_gradient(const orc_model* m, const orc_op* op, orc_ens* e, int mode, cplx* grad, cplx* E_out) {
    const int64_t P = orc_num_params(m);
    gr_ctx c;
    c.m = m; c.op = op; c.mode = mode; c.P = P; c.E = 0;
    c.O = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.EO = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.Ek = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.scratch = (cplx*)calloc((size_t)P, sizeof(cplx));
    ens_foreach(m, e, gr_fn, &c);
    const double inv = 1.0 / ens_num_steps(e);
    const cplx E = c.E * inv;
    for (int64_t k = 0; k < P; k++) {
        const cplx O = c.O[k] * inv, EO = c.EO[k] * inv, Ek = c.Ek[k] * inv;
        grad[k] = mode == 0 ? EO + conj(Ek) - 2.0 * E * conj(O) : 2.0 * EO - 2.0 * E * conj(O);
    }
    *E_out = E;
    free(c.O); free(c.EO); free(c.Ek); free(c.scratch);
}

/* ExpectationValue::fluctuation_gradient: ExpectationValue.cu:258-332 */
typedef struct {
    const orc_model* m; const orc_op* op; int64_t P;
    cplx E; double E2; cplx* O; cplx* EO; cplx* Ek; cplx* EEk; cplx* scratch;
} fg_ctx;
static void fg_fn(void* vctx, int64_t idx, const uint64_t* s, cplx lp, const cplx* theta, double w) {
    fg_ctx* c = (fg_ctx*)vctx; (void)idx;
    const cplx el = local_energy(c->m, c->op, s, lp, theta);
    c->E += w * el;
    c->E2 += w * (creal(el) * creal(el) + cimag(el) * cimag(el));
    if (c->m->kind == 0) rbm_o_k(c->m, s, theta, c->scratch); else deep_o_k(c->m, s, c->scratch);
    for (int64_t k = 2 * (int64_t)c->m->N; k < c->P; k++) {
        c->O[k] += w * c->scratch[k];
        c->EO[k] += w * el * conj(c->scratch[k]);
    }
    accumulate_E_k2(c->m, c->op, s, lp, theta, w, c->Ek, w * conj(el), c->EEk, c->scratch);
}

void orc_fluctuation_gradient(const orc_model* m, const orc_op* op, orc_ens* e, cplx* grad) {
    const int64_t P = orc_num_params(m);
    fg_ctx c;
    c.m = m; c.op = op; c.P = P; c.E = 0; c.E2 = 0;
    c.O = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.EO = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.Ek = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.EEk = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.scratch = (cplx*)calloc((size_t)P, sizeof(cplx));
    ens_foreach(m, e, fg_fn, &c);
    const double inv = 1.0 / ens_num_steps(e);
    const cplx E = c.E * inv;
    const double E2 = c.E2 * inv;
    const double nE = creal(E) * creal(E) + cimag(E) * cimag(E);
    const double fl = sqrt(E2 - nE);
    for (int64_t k = 0; k < P; k++) {
        const cplx O = c.O[k] * inv, EO = c.EO[k] * inv, Ek = c.Ek[k] * inv, EEk = c.EEk[k] * inv;
        grad[k] = 1.0 / (2.0 * fl) * (2.0 * conj(EEk - conj(E) * Ek) - 2.0 * (conj(E) * EO) + 2.0 * conj(O) * (2.0 * nE - E2));
    }
    free(c.O); free(c.EO); free(c.Ek); free(c.EEk); free(c.scratch);
}

/* ExpectationValue::difference: ExpectationValue.cu:334-446.  Samples of psi;
 * a_i = <E'_i> with E'_i the local energy of psi_prime, b_i = <r E_i> with
 * r = |psi / psi_prime|^2, result_i = a_i - b_i / <r>. */
void orc_difference(const orc_model* m, const orc_model* mp, const orc_op* const* ops, int n_ops, orc_ens* e, cplx* out) {
    const int64_t n = ens_num_steps(e);
    uint64_t* spins = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (size_t)n);
    cplx* lp = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    orc_records(m, NULL, e, spins, lp, NULL, w);
    cplx* lpp = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    orc_log_psi(mp, spins, n, lpp);
    cplx* el = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    cplx* elp = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    double ravg = 0;
    for (int64_t k = 0; k < n; k++) ravg += w[k] * exp(2.0 * (creal(lp[k]) - creal(lpp[k])));
    ravg /= (double)n;
    for (int i = 0; i < n_ops; i++) {
        orc_local_energy(m, ops[i], spins, n, el);
        orc_local_energy(mp, ops[i], spins, n, elp);
        cplx a = 0, b = 0;
        for (int64_t k = 0; k < n; k++) {
            const double r = exp(2.0 * (creal(lp[k]) - creal(lpp[k])));
            a += w[k] * elp[k];
            b += w[k] * r * el[k];
        }
        out[i] = a / (double)n - (b / (double)n) / ravg;
    }
    free(spins); free(lp); free(w); free(lpp); free(el); free(elp);
}

/* ExpectationValue::corrected: ExpectationValue.cu:74-127.  The reference
 * takes both local energies from operator_ (lines 102-106), so operator2 does
 * not enter: X = exp(2 Im E + Re(E^2 - E)); returns <E> / <X> and <X>. */
void orc_corrected(const orc_model* m, const orc_op* op, orc_ens* e, cplx* value, double* norm) {
    const int64_t n = ens_num_steps(e);
    uint64_t* spins = (uint64_t*)malloc(sizeof(uint64_t) * 2 * (size_t)n);
    cplx* el = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    orc_records(m, op, e, spins, NULL, el, w);
    cplx A = 0;
    double X = 0;
    for (int64_t k = 0; k < n; k++) {
        A += w[k] * el[k];
        X += w[k] * exp(2.0 * cimag(el[k]) + creal(el[k] * el[k] - el[k]));
    }
    A /= (double)n;
    X /= (double)n;
    *value = A / X;
    *norm = X;
    free(spins); free(el); free(w);
}

/* psi_norm: source/network_functions/PsiNorm.cu:13-44 */
typedef struct { const orc_model* m; double acc; } nm_ctx;
static void nm_fn(void* vctx, int64_t idx, const uint64_t* s, cplx lp, const cplx* theta, double w) {
    nm_ctx* c = (nm_ctx*)vctx; (void)idx; (void)s; (void)theta; (void)w;
    c->acc += probability(c->m, creal(lp));
}
double orc_psi_norm(const orc_model* m, orc_ens* e) {
    nm_ctx c = {m, 0};
    ens_foreach(m, e, nm_fn, &c);
    return sqrt(c.acc);
}

/* psi_vector: source/network_functions/PsiVector.cu:13-51 (full 2^N basis) */
void orc_psi_vector(const orc_model* m, cplx* out) {
    const int64_t n = (int64_t)1 << m->N;
    const double lpf = log(m->prefactor);
    for (int64_t idx = 0; idx < n; idx++) {
        uint64_t s[2] = {(uint64_t)idx, 0};
        mask_n(s, m->N);
        out[idx] = cexp(lpf + model_log_psi(m, s));
    }
}

/* get_S_matrix (CPU semantics): source/network_functions/S_matrix.cu:10-58.
 * Unweighted and uncentred: S[k,k'] = sum_s O_k conj(O_k') / num_steps. */
typedef struct { const orc_model* m; int64_t P; cplx* S; cplx* ok; } sm_ctx;
static void sm_fn(void* vctx, int64_t idx, const uint64_t* s, cplx lp, const cplx* theta, double w) {
    sm_ctx* c = (sm_ctx*)vctx; (void)idx; (void)lp; (void)w;
    if (c->m->kind == 0) rbm_o_k(c->m, s, theta, c->ok); else deep_o_k(c->m, s, c->ok);
    const int64_t lo = 2 * (int64_t)c->m->N;
    for (int64_t k = lo; k < c->P; k++)
        for (int64_t kp = lo; kp < c->P; kp++) c->S[k * c->P + kp] += c->ok[k] * conj(c->ok[kp]);
}
void orc_s_matrix(const orc_model* m, orc_ens* e, cplx* S) {
    const int64_t P = orc_num_params(m);
    sm_ctx c = {m, P, S, (cplx*)calloc((size_t)P, sizeof(cplx))};
    memset(S, 0, sizeof(cplx) * (size_t)(P * P));
    ens_foreach(m, e, sm_fn, &c);
    const double inv = 1.0 / ens_num_steps(e);
    for (int64_t k = 0; k < P * P; k++) S[k] *= inv;
    free(c.ok);
}

/* HilbertSpaceDistance::compute_averages: source/network_functions/
 * HilbertSpaceDistance.cu:14-136; distance 193-221; gradient 224-269.  With
 * free_quantum_axis (da != NULL) log psi' gets the first-order axis term
 * sum_i [da_i s_i r_i + i db_i (r_i cos_i - s_i sin_i)] with the single-flip
 * ratios r_i = psi'(s^i) / psi'(s), da = alpha' - alpha, db = beta' - beta,
 * cos_i / sin_i of alpha' + alpha (lines 47-80, 164-180), and the gradient
 * entries of alpha_i (index i) and beta_i (index N + i) are s_i r_i and
 * i (r_i cos_i - s_i sin_i) (lines 109-122). */
typedef struct {
    const orc_model* psi; const orc_model* psi_prime; const orc_op* op; int is_unitary; int want_grad; int64_t P;
    cplx omega_avg; double ratio_avg; double norm_avg; cplx* omega_O; cplx* ratio_O; cplx* ok;
    const double *da, *db, *ca, *sa;
    cplx* r;
} hs_ctx;
static void hs_fn(void* vctx, int64_t idx, const uint64_t* s, cplx lp, const cplx* theta, double w) {
    hs_ctx* c = (hs_ctx*)vctx; (void)idx;
    const cplx el = local_energy(c->psi, c->op, s, lp, theta);
    cplx lpp = model_log_psi(c->psi_prime, s);
    const int N = c->psi_prime->N;
    if (c->da) {
        const cplx lpp0 = lpp;
        for (int i = 0; i < N; i++) {
            uint64_t si[2] = {s[0], s[1]};
            flip(si, i);
            c->r[i] = cexp(model_log_psi(c->psi_prime, si) - lpp0);
            const double sp = spin(s, i);
            lpp += c->da[i] * sp * c->r[i] + c->db[i] * I * (c->r[i] * c->ca[i] - sp * c->sa[i]);
        }
    }
    cplx omega;
    if (c->is_unitary) {
        omega = cexp(conj(lpp - lp)) * el;
        c->norm_avg += w * creal(el * conj(el));
    } else {
        omega = cexp(el + conj(lpp - lp));
        c->norm_avg += w * exp(2 * creal(el));
    }
    const double ratio = exp(2.0 * (creal(lpp) - creal(lp)));
    c->omega_avg += w * omega;
    c->ratio_avg += w * ratio;
    if (c->want_grad) {
        orc_o_k(c->psi_prime, s, c->ok);
        for (int64_t k = 2 * (int64_t)c->psi_prime->N; k < c->P; k++) {
            c->omega_O[k] += w * omega * conj(c->ok[k]);
            c->ratio_O[k] += w * ratio * conj(c->ok[k]);
        }
        if (c->da) {
            for (int i = 0; i < N; i++) {
                const double sp = spin(s, i);
                const cplx oa = sp * c->r[i];
                const cplx ob = I * (c->r[i] * c->ca[i] - sp * c->sa[i]);
                c->omega_O[i] += w * omega * conj(oa);
                c->ratio_O[i] += w * ratio * conj(oa);
                c->omega_O[N + i] += w * omega * conj(ob);
                c->ratio_O[N + i] += w * ratio * conj(ob);
            }
        }
    }
}

double orc_hs_distance_axis(const orc_model* psi, const orc_model* psi_prime, const orc_op* op, int is_unitary,
                            orc_ens* e, int want_grad, double nu, cplx* grad, const double* alpha, const double* beta,
                            const double* alpha_prime, const double* beta_prime);

double orc_hs_distance(const orc_model* psi, const orc_model* psi_prime, const orc_op* op, int is_unitary,
                       orc_ens* e, int want_grad, double nu, cplx* grad) {
    return orc_hs_distance_axis(psi, psi_prime, op, is_unitary, e, want_grad, nu, grad, NULL, NULL, NULL, NULL);
}

double orc_hs_distance_axis(const orc_model* psi, const orc_model* psi_prime, const orc_op* op, int is_unitary,
                            orc_ens* e, int want_grad, double nu, cplx* grad, const double* alpha, const double* beta,
                            const double* alpha_prime, const double* beta_prime) {
    const int64_t P = orc_num_params(psi_prime);
    const int N = psi_prime->N;
    hs_ctx c;
    memset(&c, 0, sizeof(c));
    c.psi = psi; c.psi_prime = psi_prime; c.op = op; c.is_unitary = is_unitary; c.want_grad = want_grad; c.P = P;
    double* axis = NULL;
    if (alpha) {
        axis = (double*)malloc(sizeof(double) * 4 * (size_t)N);
        for (int i = 0; i < N; i++) {
            axis[i] = alpha_prime[i] - alpha[i];
            axis[N + i] = beta_prime[i] - beta[i];
            axis[2 * N + i] = cos(alpha_prime[i] + alpha[i]);
            axis[3 * N + i] = sin(alpha_prime[i] + alpha[i]);
        }
        c.da = axis; c.db = axis + N; c.ca = axis + 2 * N; c.sa = axis + 3 * N;
        c.r = (cplx*)calloc((size_t)N, sizeof(cplx));
    }
    c.omega_O = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.ratio_O = (cplx*)calloc((size_t)P, sizeof(cplx));
    c.ok = (cplx*)calloc((size_t)P, sizeof(cplx));
    ens_foreach(psi, e, hs_fn, &c);
    const double inv = 1.0 / ens_num_steps(e);
    const cplx om = c.omega_avg * inv;
    const double ra = c.ratio_avg * inv, na = c.norm_avg * inv;
    const double u = creal(om * conj(om));
    const double v = na * ra;
    const double distance = sqrt(1.0 - u / v);
    if (want_grad) {
        const double pref = pow(distance, nu);
        for (int64_t k = 0; k < P; k++) {
            const cplx oO = c.omega_O[k] * inv, rO = c.ratio_O[k] * inv;
            const cplx ukp = conj(om) * oO;
            const cplx vkp = na * rO;
            grad[k] = (-(ukp * v - u * vkp) / (v * v)) / pref;
        }
    }
    free(c.omega_O); free(c.ratio_O); free(c.ok); free(axis); free(c.r);
    return distance;
}
