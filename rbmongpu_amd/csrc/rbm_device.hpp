// Device-side primitives for the gfx950 VMC kernels: a minimal complex type,
// the reference's rational log-cosh approximant and its derivative, the
// Philox4x32-10 counter-based generator, spin-word helpers and 16-lane DPP
// reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RBM_DI __device__ __forceinline__
#define RBM_HDI __host__ __device__ __forceinline__

namespace rbmhip {

template <typename T>
struct cx {
    T re, im;
};

template <typename T> RBM_HDI cx<T> mk(T r, T i) { cx<T> z; z.re = r; z.im = i; return z; }
template <typename T> RBM_HDI cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <typename T> RBM_HDI cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <typename T> RBM_HDI cx<T> operator*(cx<T> a, cx<T> b) {
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T> RBM_HDI cx<T> operator*(T s, cx<T> a) { return mk(s * a.re, s * a.im); }
template <typename T> RBM_HDI cx<T> operator+(cx<T> a, T s) { return mk(a.re + s, a.im); }
template <typename T> RBM_HDI cx<T> operator-(cx<T> a, T s) { return mk(a.re - s, a.im); }
template <typename T> RBM_HDI cx<T> conj(cx<T> a) { return mk(a.re, -a.im); }
template <typename T> RBM_HDI cx<double> to_d(cx<T> a) { return mk<double>((double)a.re, (double)a.im); }

RBM_DI float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
RBM_DI double rcp(double x) { return 1.0 / x; }

// a / b computed as a * conj(b) / |b|^2
template <typename T> RBM_DI cx<T> cdiv(cx<T> a, cx<T> b) {
    const T r = rcp(b.re * b.re + b.im * b.im);
    return mk((a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r);
}

RBM_DI float fm(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
RBM_DI double fm(double a, double b, double c) { return __builtin_fma(a, b, c); }
RBM_DI float fab(float x) { return __builtin_fabsf(x); }
RBM_DI double fab(double x) { return __builtin_fabs(x); }

// exp(z) in FP32 from the hardware transcendentals: the phase is reduced to
// [-pi, pi] with a two-constant Cody-Waite step, then v_sin/v_cos take it in
// revolutions.  Measured on gfx950 against FP64: |error| <= 3.8e-7 for sin
// and cos over |Im z| <= 80, relative error of exp <= 6e-7 for |Re z| <= 8.
RBM_DI cx<float> cexp(cx<float> z) {
    const float k = __builtin_rintf(z.im * 0.15915494309189535f);
    float r = __builtin_fmaf(-k, 6.28318548202514648f, z.im);
    r = __builtin_fmaf(-k, -1.7484555e-07f, r);
    const float rev = r * 0.15915494309189535f;
    const float e = __expf(z.re);
    return mk(e * __builtin_amdgcn_cosf(rev), e * __builtin_amdgcn_sinf(rev));
}
// exp(z) delivered in FP64 with FP64's exponent range: exp(Re z) = 2^k 2^f
// with k = floor(Re z log2 e), 2^f from v_exp_f32 and 2^k applied by
// v_ldexp_f64, so a ratio psi(s')/psi(s) beyond FP32's range (Re z > 88)
// stays finite.  Relative accuracy as cexp above.
RBM_DI cx<double> cexp_wide(cx<float> z) {
    const float t = z.re * 1.4426950408889634f;
    const float k = __builtin_floorf(t);
    const float m = __builtin_amdgcn_exp2f(t - k);
    const float kk = __builtin_rintf(z.im * 0.15915494309189535f);
    float r = __builtin_fmaf(-kk, 6.28318548202514648f, z.im);
    r = __builtin_fmaf(-kk, -1.7484555e-07f, r);
    const float rev = r * 0.15915494309189535f;
    const int e = (int)__builtin_fmaxf(__builtin_fminf(k, 2000.0f), -2000.0f);
    return mk(__builtin_ldexp((double)(m * __builtin_amdgcn_cosf(rev)), e),
              __builtin_ldexp((double)(m * __builtin_amdgcn_sinf(rev)), e));
}
RBM_DI cx<double> cexp(cx<double> z) {
    double s, c;
    sincos(z.im, &s, &c);
    const double e = exp(z.re);
    return mk(e * c, e * s);
}
RBM_DI cx<double> cexp_wide(cx<double> z) { return cexp(z); }

// Shared denominator of the approximant:
// 9.19376335670885 + z (s 10.218.. + z (7.7714.. + z (s 3.7466.. + z)))
// (include/quantum_state/psi_functions.hpp:56, 103).
template <typename T> RBM_DI cx<T> lc_den(cx<T> z, T sg, T c1) {
    cx<T> d = z + sg * T(3.746646023906276);
    d = z * d + T(7.771429504240965);
    d = z * d + sg * c1;
    d = z * d + T(9.19376335670885);
    return d;
}

// my_logcosh (include/quantum_state/psi_functions.hpp:49-57):
//   lc(z) = sg k z + (n0 - sg n1 z) / D_sg(z) - c,   sg = Re z > 0 ? 1 : -1,
//   D_sg(z) = z^4 + sg a3 z^3 + a2 z^2 + sg a1 z + a0.
// Every odd power of z carries one factor sg, so with w = sg z the function is
// L(w) = k w + (n0 - n1 w) / D_+(w) - c.  L has real coefficients, hence
// L(conj w) = conj L(w); with v = |Re z| + i Im z (w = v for sg = 1 and
// w = conj v for sg = -1):
//   Re lc(z) = Re L(v),   Im lc(z) = sg Im L(v).
// This evaluates the reference's function (its branch at Re z = 0 included)
// without sign-dependent coefficients.
constexpr double LC_A3 = 3.746646023906276, LC_A2 = 7.771429504240965, LC_A1 = 10.2180213465,
                 LC_A0 = 9.19376335670885, LC_N0 = 5.49914721954, LC_N1 = 2.16564366435,
                 LC_K = 0.9003320053750442, LC_C = 0.598139;

// D_+(w) = (((w + a3) w + a2) w + a1) w + a0 at w = wr + i wi.
// FP32: D_+ has real coefficients and w is a root of x^2 - s x + t with
// s = 2 wr, t = |w|^2, so D_+(w) is the remainder R1 w + R0 of D_+ divided by
// that quadratic: with c3 = a3 + s, c2 = a2 - t + s c3,
// R1 = a1 + s c2 - t c3 and R0 = a0 - t c2 (11 instructions against 13 for
// the complex Horner form; the same accuracy for |w| below 1 and within 2x
// of it up to |w| ~ 10).  FP64 keeps the Horner form.
template <typename T> RBM_DI void lc_den_w(T wr, T wi, T& dr, T& di) {
    if constexpr (sizeof(T) == 4) {
        const T s = wr + wr;
        const T t = fm(wr, wr, wi * wi);
        const T c3 = s + T(LC_A3);
        const T c2 = fm(s, c3, T(LC_A2) - t);
        const T r1 = fm(s, c2, fm(-t, c3, T(LC_A1)));
        const T r0 = fm(-t, c2, T(LC_A0));
        dr = fm(r1, wr, r0);
        di = r1 * wi;
        return;
    }
    const T ur = wr + T(LC_A3);
    const T vr = fm(wr, ur, fm(-wi, wi, T(LC_A2)));
    const T vi = wi * (wr + ur);
    const T xr = fm(wr, vr, fm(-wi, vi, T(LC_A1)));
    const T xi = fm(wr, vi, wi * vr);
    dr = fm(wr, xr, fm(-wi, xi, T(LC_A0)));
    di = fm(wr, xi, wi * xr);
}

template <typename T> RBM_DI cx<T> logcosh(cx<T> z) {
    const T wr = fab(z.re), wi = z.im;
    T dr, di;
    lc_den_w(wr, wi, dr, di);
    const T nr = fm(T(-LC_N1), wr, T(LC_N0));
    const T ni = T(-LC_N1) * wi;
    const T r = rcp(fm(dr, dr, di * di));
    const T qr = fm(nr, dr, ni * di);
    const T qi = fm(ni, dr, -(nr * di));
    const T lr = fm(qr, r, fm(T(LC_K), wr, T(-LC_C)));
    const T li = fm(qi, r, T(LC_K) * wi);
    return mk(lr, z.re > T(0) ? li : -li);
}

// my_logcosh(z) - L0 without cancellation, L0 = n0/a0 - c = my_logcosh(0).
// The direct form adds k w, n0 / a0 + O(w) and -c, which cancel down to
// about w^2 / 2: at |z| ~ 0.1 an FP32 evaluation keeps only ~3e-8 absolute.
// Over the common denominator the constant and linear parts cancel in the
// coefficients instead (computed in FP64 from the reference's constants):
//   my_logcosh(z) = L0 + w (e0 + e1 w + e2 w^2 + e3 w^3 + e4 w^4) / D_+(w),
//   e_i = (k a0 [a0, a1, a2, a3, 1]_i - [a0 n1 + n0 a1, n0 a2, n0 a3, n0, 0]_i) / a0,
// with e0 ~ 3e-11 and e1 ~ 4.55, so a small w keeps its relative accuracy.
// Used where a sum of my_logcosh values is itself the result (log psi).
constexpr double LC_L0 = LC_N0 / LC_A0 - LC_C;
constexpr double LC_E0 = (LC_K * LC_A0 * LC_A0 - (LC_A0 * LC_N1 + LC_N0 * LC_A1)) / LC_A0;
constexpr double LC_E1 = (LC_K * LC_A0 * LC_A1 - LC_N0 * LC_A2) / LC_A0;
constexpr double LC_E2 = (LC_K * LC_A0 * LC_A2 - LC_N0 * LC_A3) / LC_A0;
constexpr double LC_E3 = (LC_K * LC_A0 * LC_A3 - LC_N0) / LC_A0;
constexpr double LC_E4 = LC_K;

template <typename T> RBM_DI cx<T> logcosh_shifted(cx<T> z) {
    const cx<T> v = mk(fab(z.re), z.im);
    T dr, di;
    lc_den_w(v.re, v.im, dr, di);
    cx<T> p = T(LC_E4) * v + mk(T(LC_E3), T(0));
    p = v * p + mk(T(LC_E2), T(0));
    p = v * p + mk(T(LC_E1), T(0));
    p = v * p + mk(T(LC_E0), T(0));
    const cx<T> q = cdiv(v * p, mk(dr, di));
    return mk(q.re, z.re > T(0) ? q.im : -q.im);
}

// Real part of my_logcosh only (the Metropolis test needs Re log psi).
template <typename T> RBM_DI T logcosh_re(cx<T> z) {
    const T wr = fab(z.re), wi = z.im;
    T dr, di;
    lc_den_w(wr, wi, dr, di);
    const T nr = fm(T(-LC_N1), wr, T(LC_N0));
    const T num = fm(nr, dr, T(-LC_N1) * (wi * di));
    const T r = rcp(fm(dr, dr, di * di));
    return fm(num, r, fm(T(LC_K), wr, T(-LC_C)));
}

// my_logcosh(z + h) - my_logcosh(z) without subtracting two evaluations.
// When z and z + h lie on the same side of Re = 0, w = sg z, w' = w + sg h and
// L(w') - L(w) = sg h [k - (n0 Q + n1 R) / (D(w) D(w'))] with the divided
// differences of the rational part, written in s = w + w', p = w w':
//   Q = (D(w') - D(w)) / (w' - w) = s (s^2 - 2p) + a3 (s^2 - p) + a2 s + a1,
//   R = (w' D(w) - w D(w')) / (w' - w) = a0 - p (a2 + a3 s + s^2 - p),
// so a small h keeps its relative accuracy.  Across Re = 0 (the reference's
// branch jump) the two sides are evaluated separately.
template <typename T> RBM_DI cx<T> logcosh_diff(cx<T> z, cx<T> h) {
    const cx<T> z2 = z + h;
    const bool up = z.re > T(0);
    if (up != (z2.re > T(0))) return logcosh(z2) - logcosh(z);
    const T sg = up ? T(1) : T(-1);
    const cx<T> w = sg * z, w2 = sg * z2, hh = sg * h;
    const cx<T> d1 = lc_den(w, T(1), T(LC_A1)), d2 = lc_den(w2, T(1), T(LC_A1));
    const cx<T> sm = w + w2, pr = w * w2;
    const cx<T> s2 = sm * sm;
    const cx<T> q = sm * (s2 - T(2) * pr) + T(LC_A3) * (s2 - pr) + T(LC_A2) * sm + mk(T(LC_A1), T(0));
    const cx<T> r = mk(T(LC_A0), T(0)) - pr * ((s2 - pr) + T(LC_A3) * sm + mk(T(LC_A2), T(0)));
    const cx<T> num = T(LC_N0) * q + T(LC_N1) * r;
    return hh * (mk(T(LC_K), T(0)) - cdiv(num, d1 * d2));
}

// my_tanh: include/quantum_state/psi_functions.hpp:97-114.
template <typename T> RBM_DI cx<T> dlogcosh(cx<T> z) {
    const T sg = z.re > T(0) ? T(1) : T(-1);
    const cx<T> d = lc_den(z, sg, T(10.218021346543315));
    cx<T> p = mk(T(6.746450656267947) + sg * T(0.9003320053750442) * z.re, sg * T(0.9003320053750442) * z.im);
    p = z * p + sg * T(26.632014683761202);
    p = z * p + T(70.82878897882324);
    p = z * p + sg * T(146.36284300074402);
    p = z * p + T(199.24474920889975);
    p = z * p + sg * T(177.6769746361748);
    p = z * p + T(83.68563506532087);
    p = z * p;
    return cdiv(p, d * d);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  One call yields 128 bits.
// ---------------------------------------------------------------------------
struct u32x4 { uint32_t x, y, z, w; };

RBM_HDI u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1; c3 = (uint32_t)p0; c0 = n0; c2 = n2;
    }
    u32x4 r; r.x = c0; r.y = c1; r.z = c2; r.w = c3;
    return r;
}

// ---------------------------------------------------------------------------
// Spin words: spin i = +1 iff bit i is set (include/Spins.h:125-129).  Two
// 64-bit words cover N <= 128.
// ---------------------------------------------------------------------------
RBM_HDI int bit_of(uint64_t s0, uint64_t s1, int i) {
    return (int)(((i < 64) ? (s0 >> i) : (s1 >> (i - 64))) & 1u);
}
template <typename T> RBM_HDI T spin_of(uint64_t s0, uint64_t s1, int i) {
    return bit_of(s0, s1, i) ? T(1) : T(-1);
}
RBM_HDI void flip_bit(uint64_t& s0, uint64_t& s1, int i) {
    if (i < 64) s0 ^= (uint64_t)1 << i; else s1 ^= (uint64_t)1 << (i - 64);
}
RBM_HDI void mask_words(uint64_t& s0, uint64_t& s1, int N) {
    if (N < 64) { s0 &= ((uint64_t)1 << N) - 1u; s1 = 0; }
    else if (N == 64) { s1 = 0; }
    else if (N < 128) { s1 &= ((uint64_t)1 << (N - 64)) - 1u; }
}

// (-1)^popcount(zy & ~s): the product of s_i over the bits of zy.
RBM_HDI double zy_sign(uint64_t z0, uint64_t z1, uint64_t s0, uint64_t s1) {
    const int par = (__builtin_popcountll(z0 & ~s0) + __builtin_popcountll(z1 & ~s1)) & 1;
    return par ? -1.0 : 1.0;
}

// ---------------------------------------------------------------------------
// 16-lane reductions with DPP row rotations.  Every lane of the row ends with
// the same bits: at each level a lane adds its partner's partial, and the two
// partners add the same two operands (the sum is commutative), so the whole
// row agrees on the accept/reject decision.
// ---------------------------------------------------------------------------
template <int CTRL> RBM_DI float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> RBM_DI double dpp(double v) {
    const int2 w = *reinterpret_cast<int2*>(&v);
    int2 r;
    r.x = __builtin_amdgcn_update_dpp(0, w.x, CTRL, 0xF, 0xF, false);
    r.y = __builtin_amdgcn_update_dpp(0, w.y, CTRL, 0xF, 0xF, false);
    return *reinterpret_cast<double*>(&r);
}

// row_ror:8, 4, 2, 1 (dpp_ctrl 0x120 + n)
template <typename T> RBM_DI T row_sum16(T v) {
    v += dpp<0x128>(v);
    v += dpp<0x124>(v);
    v += dpp<0x122>(v);
    v += dpp<0x121>(v);
    return v;
}
template <typename T> RBM_DI cx<T> row_sum16(cx<T> v) {
    v.re += dpp<0x128>(v.re); v.im += dpp<0x128>(v.im);
    v.re += dpp<0x124>(v.re); v.im += dpp<0x124>(v.im);
    v.re += dpp<0x122>(v.re); v.im += dpp<0x122>(v.im);
    v.re += dpp<0x121>(v.re); v.im += dpp<0x121>(v.im);
    return v;
}

// ---------------------------------------------------------------------------
// Operator tables after host canonicalisation.  A Pauli string becomes
// (flip mask, coefficient * (-i)^{#Y}, zy mask): its matrix element at s is
// coef * prod_{i in zy} s_i and it maps s to s ^ mask
// (include/operator/Operator.hpp:45-88).  Strings are grouped by flip mask;
// mask 0 is the diagonal part.
// ---------------------------------------------------------------------------
struct OpDev {
    int n_diag;
    const cx<double>* diag_coef;   // n_diag
    const uint64_t* diag_zy;       // 2 * n_diag
    int n_terms;                   // off-diagonal flip masks
    const int* flip_begin;         // n_terms + 1
    const int* flips;              // flip positions of each term
    const int* ent_begin;          // n_terms + 1
    const cx<double>* ent_coef;    // entries
    const uint64_t* ent_zy;        // 2 * entries
};

// Operator tables are read-only for the lifetime of a launch: reading them
// through the constant address space lets wave-uniform indices become scalar
// loads.
#define RBM_CONST __attribute__((address_space(4)))
template <typename T> RBM_DI T ldc(const T* p, int i) { return ((const RBM_CONST T*)p)[i]; }
RBM_DI cx<double> ldc(const cx<double>* p, int i) {
    const RBM_CONST double* q = (const RBM_CONST double*)p;
    return mk(q[2 * i], q[2 * i + 1]);
}

RBM_DI cx<double> diag_energy(const OpDev& op, uint64_t s0, uint64_t s1) {
    cx<double> e = mk(0.0, 0.0);
    for (int d = 0; d < op.n_diag; d++) {
        const double sg = zy_sign(ldc(op.diag_zy, 2 * d), ldc(op.diag_zy, 2 * d + 1), s0, s1);
        e = e + sg * ldc(op.diag_coef, d);
    }
    return e;
}

RBM_DI cx<double> term_coef(const OpDev& op, int t, uint64_t s0, uint64_t s1) {
    cx<double> c = mk(0.0, 0.0);
    const int ee = ldc(op.ent_begin, t + 1);
    for (int e = ldc(op.ent_begin, t); e < ee; e++) {
        const uint64_t z0 = ldc(op.ent_zy, 2 * e), z1 = ldc(op.ent_zy, 2 * e + 1);
        const cx<double> ce = ldc(op.ent_coef, e);
        // entries without Z or Y factors (uniform test) need no parity
        c = (z0 | z1) ? c + zy_sign(z0, z1, s0, s1) * ce : c + ce;
    }
    return c;
}

// ---------------------------------------------------------------------------
// Move types of MonteCarloLoop other than the single flip
// (include/spin_ensembles/MonteCarloLoop.hpp:79-93, 153-201).  Proposal k of
// chain c in call `call` draws from Philox(key = seed, ctr = (c, k, call, .)):
//   exchange (set_total_z_symmetry): block 4 = (x, y, z, .): flip p = x % N, then
//     flip the (y % cnt)-th (in index order) of the cnt spins equal to the new
//     value of p (p itself included, which undoes the move); u from z.  The
//     reference draws the second spin by rejection, uniformly over the same set.
//   fast sweep (set_fast_sweep): flip mask = AND of max(tries, 1) random N-bit
//     strings, blocks 0x10 + j; u from block 5, word x.
// A chain in a sector starts from its random configuration (block 1 of k = 0)
// with single bits set or cleared until k_up spins are up, bit j chosen by
// block (c, j, call, 6).
// ---------------------------------------------------------------------------
RBM_HDI int select_bit(uint64_t w0, uint64_t w1, int k) {
    const int c0 = __builtin_popcountll(w0);
    uint64_t w = w0;
    int base = 0;
    if (k >= c0) { k -= c0; w = w1; base = 64; }
    for (int i = 0; i < k; i++) w &= w - 1u;
    return base + __builtin_ctzll(w);
}

RBM_HDI void full_mask(int N, uint64_t& m0, uint64_t& m1) {
    m0 = ~(uint64_t)0; m1 = ~(uint64_t)0;
    mask_words(m0, m1, N);
}

RBM_HDI void sector_start(int N, int k_up, uint32_t chain, uint32_t call, uint32_t sk0, uint32_t sk1,
                          uint64_t& s0, uint64_t& s1) {
    uint64_t m0, m1;
    full_mask(N, m0, m1);
    for (uint32_t j = 0;; j++) {
        const int up = __builtin_popcountll(s0) + __builtin_popcountll(s1);
        if (up == k_up) break;
        const u32x4 r = philox(chain, j, call, 6u, sk0, sk1);
        if (up > k_up) {
            flip_bit(s0, s1, select_bit(s0, s1, (int)(r.x % (uint32_t)up)));
        } else {
            flip_bit(s0, s1, select_bit(~s0 & m0, ~s1 & m1, (int)(r.x % (uint32_t)(N - up))));
        }
    }
}

// The proposed configuration (n0, n1) and the uniform's 32-bit word.
RBM_HDI void propose_move(int move, int tries, int N, uint32_t chain, uint64_t k, uint32_t call, uint32_t sk0,
                          uint32_t sk1, uint64_t s0, uint64_t s1, uint64_t& n0, uint64_t& n1, uint32_t& uword) {
    uint64_t m0, m1;
    full_mask(N, m0, m1);
    if (move == 1) {
        const u32x4 r = philox(chain, (uint32_t)k, call, 4u, sk0, sk1);
        const int p = (int)(r.x % (uint32_t)N);
        n0 = s0; n1 = s1;
        flip_bit(n0, n1, p);
        const bool v = bit_of(n0, n1, p) != 0;
        const uint64_t c0 = v ? n0 : (~n0 & m0), c1 = v ? n1 : (~n1 & m1);
        const int cnt = __builtin_popcountll(c0) + __builtin_popcountll(c1);
        flip_bit(n0, n1, select_bit(c0, c1, (int)(r.y % (uint32_t)cnt)));
        uword = r.z;
    } else {
        uint64_t f0 = m0, f1 = m1;
        const int nt = tries > 1 ? tries : 1;
        for (int j = 0; j < nt; j++) {
            const u32x4 r = philox(chain, (uint32_t)k, call, 0x10u + (uint32_t)j, sk0, sk1);
            f0 &= (uint64_t)r.x | ((uint64_t)r.y << 32);
            f1 &= (uint64_t)r.z | ((uint64_t)r.w << 32);
        }
        n0 = s0 ^ f0; n1 = s1 ^ f1;
        uword = philox(chain, (uint32_t)k, call, 5u, sk0, sk1).x;
    }
}

// The same matrix element for a term index that differs between lanes (plain
// vector loads instead of scalar ones).
RBM_DI cx<double> term_coef_lane(const OpDev& op, int t, uint64_t s0, uint64_t s1) {
    cx<double> c = mk(0.0, 0.0);
    for (int e = op.ent_begin[t]; e < op.ent_begin[t + 1]; e++) {
        const uint64_t z0 = op.ent_zy[2 * e], z1 = op.ent_zy[2 * e + 1];
        c = c + zy_sign(z0, z1, s0, s1) * op.ent_coef[e];
    }
    return c;
}

}  // namespace rbmhip
