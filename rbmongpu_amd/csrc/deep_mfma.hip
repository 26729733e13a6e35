// PsiDeep on BF16 MFMA (gfx950): log psi of 16 configurations per wavefront
// with every layer as a dense GEMM on v_mfma_f32_16x16x32_bf16, FP32
// accumulation and FP32 my_logcosh activations (include/quantum_state/
// PsiDeep.hpp:89-133 forward_pass, 135-169 log_psi_s_generic without
// translational invariance).
//
// Layer l computes A_l^T (units x inputs) times X^T (inputs x configs).  The
// B operand holds the 16 configurations in its columns, so a lane's four
// accumulator rows of one 16-unit output tile are exactly the B-operand rows
// that lane supplies for the next layer: two consecutive output tiles form one
// 32-wide K block of the next layer without any data exchange between lanes.
// The weight fragments are stored in that K order by the host.
//
// A complex layer a = W^T x + b is four real products:
//   Re a = Wr^T xr + Wi^T (-xi),  Im a = Wi^T xr + Wr^T xi
// (two for the first layer, whose input spins are real).
//
// The weights (bfloat16 fragments, 1 KiB per 16 x 32 block and part) and the
// FP32 biases and final weights live in LDS for the whole kernel.
#include "deep_kernels.hpp"

namespace rbmhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

RBM_DI bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

RBM_DI uint32_t pack_bf16(float lo, float hi) {
    const __bf16 a = (__bf16)lo, b = (__bf16)hi;
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}

// +-1 as bfloat16 for spin i (bit set = +1), 0 beyond N.
RBM_DI uint32_t spin_bf16(uint32_t bits, int q, int i, int N) {
    const uint32_t lo = i < N ? (((bits >> (2 * q)) & 1u) ? 0x3F80u : 0xBF80u) : 0u;
    const uint32_t hi = i + 1 < N ? (((bits >> (2 * q + 1)) & 1u) ? 0x3F80u : 0xBF80u) : 0u;
    return lo | (hi << 16);
}

// Complex log psi (without the prefactor) of the configuration in this lane's
// column; every lane of the column returns the same bits.  F: weight
// fragments, V: biases and final weights, both in LDS.
template <int KBMAX, bool REC = false>
RBM_DI cx<float> mfma_log_psi(const DeepMfmaDev& md, const uint4* __restrict__ F, const float* __restrict__ V,
                              uint64_t s0, uint64_t s1, int lane, cx<float>* row = nullptr) {
    const int g = lane >> 4;
    uint32_t xr[KBMAX][4], xi[KBMAX][4];
#pragma unroll
    for (int kb = 0; kb < KBMAX; kb++) {
        const int i0 = kb * 32 + 8 * g;
        const uint32_t bits = i0 < 64 ? (uint32_t)(s0 >> i0) & 0xFFu : (i0 < 128 ? (uint32_t)(s1 >> (i0 - 64)) & 0xFFu : 0u);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            xr[kb][q] = spin_bf16(bits, q, i0 + 2 * q, md.N);
            xi[kb][q] = 0u;
        }
    }
    cx<float> part = mk(0.0f, 0.0f);
    for (int l = 0; l < md.L; l++) {
        const int MT = md.out_p[l] >> 4, KB = md.in_p[l] >> 5;
        const uint4* Fl = F + md.frag_off[l];
        const float* br = V + md.vec_off[l];
        const float* bi = br + md.out_p[l];
        const bool last = l == md.L - 1;
        uint32_t nr[KBMAX][4], ni[KBMAX][4];
#pragma unroll
        for (int mt = 0; mt < 2 * KBMAX; mt++) {
            if (mt < MT) {
                const int u0 = mt * 16 + 4 * g;
                f32x4 ar = *reinterpret_cast<const f32x4*>(br + u0);
                f32x4 ai = *reinterpret_cast<const f32x4*>(bi + u0);
#pragma unroll
                for (int kb = 0; kb < KBMAX; kb++) {
                    if (kb < KB) {
                        const uint4* fb = Fl + (size_t)((mt * KB + kb) * 2) * 64 + lane;
                        const bf16x8 Ar = as_bf16x8(fb[0]);
                        const bf16x8 Ai = as_bf16x8(fb[64]);
                        const bf16x8 Br = as_bf16x8(make_uint4(xr[kb][0], xr[kb][1], xr[kb][2], xr[kb][3]));
                        ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Br, ar, 0, 0, 0);
                        ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, Br, ai, 0, 0, 0);
                        if (l > 0) {
                            const bf16x8 Bi = as_bf16x8(make_uint4(xi[kb][0], xi[kb][1], xi[kb][2], xi[kb][3]));
                            const bf16x8 Bn = as_bf16x8(make_uint4(xi[kb][0] ^ 0x80008000u, xi[kb][1] ^ 0x80008000u,
                                                                   xi[kb][2] ^ 0x80008000u, xi[kb][3] ^ 0x80008000u));
                            ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, Bn, ar, 0, 0, 0);
                            ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Bi, ai, 0, 0, 0);
                        }
                    }
                }
                cx<float> x[4];
#pragma unroll
                for (int v = 0; v < 4; v++) x[v] = logcosh(mk(ar[v], ai[v]));
                if (REC && row) {
                    // x_l and, in the delta slots until the backward pass, my_tanh(a_l)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        row[md.rec_x[l] + u0 + v] = x[v];
                        row[md.rec_d[l] + u0 + v] = dlogcosh(mk(ar[v], ai[v]));
                    }
                }
                if (last) {
                    const float* fr = V + md.final_off;
                    const float* fi = fr + md.out_p[l];
                    const f32x4 wr = *reinterpret_cast<const f32x4*>(fr + u0);
                    const f32x4 wi = *reinterpret_cast<const f32x4*>(fi + u0);
#pragma unroll
                    for (int v = 0; v < 4; v++) part = part + x[v] * mk(wr[v], wi[v]);
                } else {
                    const int q = (mt & 1) * 2;
                    nr[mt >> 1][q] = pack_bf16(x[0].re, x[1].re);
                    nr[mt >> 1][q + 1] = pack_bf16(x[2].re, x[3].re);
                    ni[mt >> 1][q] = pack_bf16(x[0].im, x[1].im);
                    ni[mt >> 1][q + 1] = pack_bf16(x[2].im, x[3].im);
                }
            }
        }
        if (!last) {
#pragma unroll
            for (int kb = 0; kb < KBMAX; kb++)
#pragma unroll
                for (int q = 0; q < 4; q++) { xr[kb][q] = nr[kb][q]; xi[kb][q] = ni[kb][q]; }
        }
    }
    // sum over the four lane groups of the column (butterfly: equal bits)
    part.re += __shfl_xor(part.re, 16);
    part.im += __shfl_xor(part.im, 16);
    part.re += __shfl_xor(part.re, 32);
    part.im += __shfl_xor(part.im, 32);
    return part;
}

// Forward pass of a two-layer network whose padded shape is known at compile
// time: layer 0 has MT0 output tiles over KB0 K blocks, layer 1 MT1 tiles over
// KB1 = MT0 / 2 K blocks, and K block k of layer 1 is made of layer-0 tiles
// 2k and 2k + 1.  The pass is software-pipelined over pairs of layer-0 tiles:
// step k issues the layer-0 products of pair k + 1 and the layer-1 products of
// K block k - 1, then evaluates my_logcosh of pair k, so matrix-core work is
// always in flight under the vector work; a scheduling barrier closes each
// step so that fragment loads are not all hoisted.  Same arithmetic, and the
// same order of every accumulation, as mfma_log_psi.
template <int MT0, int KB0, int MT1>
RBM_DI cx<float> mfma_log_psi_2l(const DeepMfmaDev& md, const uint4* __restrict__ F, const float* __restrict__ V,
                                 uint64_t s0, uint64_t s1, int lane) {
    static_assert(MT0 % 2 == 0, "pairs of layer-0 tiles form layer-1 K blocks");
    constexpr int NP = MT0 / 2;   // pairs = layer-1 K blocks
    const int g = lane >> 4;
    uint32_t xs[KB0][4];
#pragma unroll
    for (int kb = 0; kb < KB0; kb++) {
        const int i0 = kb * 32 + 8 * g;
        const uint32_t bits = i0 < 64 ? (uint32_t)(s0 >> i0) & 0xFFu : (i0 < 128 ? (uint32_t)(s1 >> (i0 - 64)) & 0xFFu : 0u);
#pragma unroll
        for (int q = 0; q < 4; q++) xs[kb][q] = spin_bf16(bits, q, i0 + 2 * q, md.N);
    }
    const uint4* F0 = F + md.frag_off[0];
    const uint4* F1 = F + md.frag_off[1];
    const float* b0r = V + md.vec_off[0];
    const float* b0i = b0r + md.out_p[0];
    const float* b1r = V + md.vec_off[1];
    const float* b1i = b1r + md.out_p[1];
    // layer-0 products of tiles 2p, 2p + 1 into (ar, ai)
    auto layer0_pair = [&](int p, f32x4 (&ar)[2], f32x4 (&ai)[2]) {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            ar[i] = *reinterpret_cast<const f32x4*>(b0r + (2 * p + i) * 16 + 4 * g);
            ai[i] = *reinterpret_cast<const f32x4*>(b0i + (2 * p + i) * 16 + 4 * g);
        }
#pragma unroll
        for (int kb = 0; kb < KB0; kb++) {
            const bf16x8 B = as_bf16x8(make_uint4(xs[kb][0], xs[kb][1], xs[kb][2], xs[kb][3]));
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const uint4* fb = F0 + (size_t)(((2 * p + i) * KB0 + kb) * 2) * 64 + lane;
                ar[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(fb[0]), B, ar[i], 0, 0, 0);
                ai[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(fb[64]), B, ai[i], 0, 0, 0);
            }
        }
    };
    f32x4 yr[MT1], yi[MT1];   // layer-1 accumulators
#pragma unroll
    for (int mt = 0; mt < MT1; mt++) {
        yr[mt] = *reinterpret_cast<const f32x4*>(b1r + mt * 16 + 4 * g);
        yi[mt] = *reinterpret_cast<const f32x4*>(b1i + mt * 16 + 4 * g);
    }
    // layer-1 products of K block kb with its packed inputs
    auto layer1_block = [&](int kb, const uint32_t (&xr)[4], const uint32_t (&xi)[4]) {
        const bf16x8 Br = as_bf16x8(make_uint4(xr[0], xr[1], xr[2], xr[3]));
        const bf16x8 Bi = as_bf16x8(make_uint4(xi[0], xi[1], xi[2], xi[3]));
        const bf16x8 Bn = as_bf16x8(make_uint4(xi[0] ^ 0x80008000u, xi[1] ^ 0x80008000u, xi[2] ^ 0x80008000u,
                                               xi[3] ^ 0x80008000u));
#pragma unroll
        for (int mt = 0; mt < MT1; mt++) {
            const uint4* fb = F1 + (size_t)((mt * NP + kb) * 2) * 64 + lane;
            const bf16x8 Ar = as_bf16x8(fb[0]);
            const bf16x8 Ai = as_bf16x8(fb[64]);
            yr[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Br, yr[mt], 0, 0, 0);
            yi[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, Br, yi[mt], 0, 0, 0);
            yr[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, Bn, yr[mt], 0, 0, 0);
            yi[mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Bi, yi[mt], 0, 0, 0);
        }
    };
    f32x4 cr[2], ci[2];
    uint32_t pr[4], pi[4];   // packed activations of the previous pair (K block p - 1)
    layer0_pair(0, cr, ci);
#pragma unroll
    for (int p = 0; p < NP; p++) {
        f32x4 nr[2], ni[2];
        if (p + 1 < NP) layer0_pair(p + 1, nr, ni);
        if (p > 0) layer1_block(p - 1, pr, pi);
#pragma unroll
        for (int i = 0; i < 2; i++) {
            cx<float> x[4];
#pragma unroll
            for (int v = 0; v < 4; v++) x[v] = logcosh(mk(cr[i][v], ci[i][v]));
            pr[2 * i] = pack_bf16(x[0].re, x[1].re);
            pr[2 * i + 1] = pack_bf16(x[2].re, x[3].re);
            pi[2 * i] = pack_bf16(x[0].im, x[1].im);
            pi[2 * i + 1] = pack_bf16(x[2].im, x[3].im);
        }
        if (p + 1 < NP) {
#pragma unroll
            for (int i = 0; i < 2; i++) { cr[i] = nr[i]; ci[i] = ni[i]; }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    layer1_block(NP - 1, pr, pi);
    cx<float> part = mk(0.0f, 0.0f);
    const float* fr = V + md.final_off;
    const float* fi = fr + md.out_p[1];
#pragma unroll
    for (int mt = 0; mt < MT1; mt++) {
        const int u0 = mt * 16 + 4 * g;
        const f32x4 wr = *reinterpret_cast<const f32x4*>(fr + u0);
        const f32x4 wi = *reinterpret_cast<const f32x4*>(fi + u0);
#pragma unroll
        for (int v = 0; v < 4; v++) part = part + logcosh(mk(yr[mt][v], yi[mt][v])) * mk(wr[v], wi[v]);
    }
    part.re += __shfl_xor(part.re, 16);
    part.im += __shfl_xor(part.im, 16);
    part.re += __shfl_xor(part.re, 32);
    part.im += __shfl_xor(part.im, 32);
    return part;
}

// The forward pass a kernel instantiation uses: any shape within KBMAX blocks,
// or one compile-time two-layer shape.
template <int KBMAX>
struct FwdAny {
    RBM_DI static cx<float> run(const DeepMfmaDev& md, const uint4* F, const float* V, uint64_t s0, uint64_t s1, int lane) {
        return mfma_log_psi<KBMAX>(md, F, V, s0, s1, lane);
    }
};
template <int MT0, int KB0, int MT1>
struct Fwd2 {
    RBM_DI static cx<float> run(const DeepMfmaDev& md, const uint4* F, const float* V, uint64_t s0, uint64_t s1, int lane) {
        return mfma_log_psi_2l<MT0, KB0, MT1>(md, F, V, s0, s1, lane);
    }
};
// the deep-network tier, 40 -> 160 -> 80: in_p 64, 160 and out_p 160, 80
using FwdDense40x160x80 = Fwd2<10, 2, 5>;

static inline bool is_shape_40x160x80(const DeepMfmaDev& md) {
    return md.L == 2 && md.in_p[0] == 64 && md.out_p[0] == 160 && md.in_p[1] == 160 && md.out_p[1] == 80;
}

// Copy the vectors and the first nfrag fragments into LDS (dynamic shared
// memory): vectors first, then the forward fragments, then the backward ones.
RBM_DI void mfma_stage(const DeepMfmaDev& md, unsigned char* smem, int nfrag, const uint4** F, const float** V) {
    float* v = reinterpret_cast<float*>(smem);
    for (int i = threadIdx.x; i < md.vec_total; i += blockDim.x) v[i] = md.vec[i];
    uint4* f = reinterpret_cast<uint4*>(v + md.vec_total);
    for (int i = threadIdx.x; i < nfrag; i += blockDim.x) f[i] = md.frag[i];
    __syncthreads();
    *F = f;
    *V = v;
}

// Back-propagation (PsiDeep.hpp:214-293) for the column's configuration after
// mfma_log_psi<KBMAX, true>: delta_{L-1} = final * my_tanh(a_{L-1}), then
// delta_l = (W_{l+1} delta_{l+1}) * my_tanh(a_l) on the backward fragments;
// every lane reads back only the entries it wrote in the forward pass and
// replaces my_tanh(a_l) by delta_l.  delta_{l+1} enters the MFMA as bfloat16.
template <int KBMAX>
RBM_DI void mfma_backward(const DeepMfmaDev& md, const uint4* __restrict__ F, const float* __restrict__ V,
                          cx<float>* row, int lane) {
    const int g = lane >> 4;
    uint32_t dr[KBMAX][4], di[KBMAX][4];
#pragma unroll
    for (int kb = 0; kb < KBMAX; kb++)
#pragma unroll
        for (int q = 0; q < 4; q++) { dr[kb][q] = 0u; di[kb][q] = 0u; }
    {
        const int l = md.L - 1, MT = md.out_p[l] >> 4;
        const float* fr = V + md.final_off;
        const float* fi = fr + md.out_p[l];
#pragma unroll
        for (int mt = 0; mt < 2 * KBMAX; mt++) {
            if (mt < MT) {
                const int u0 = mt * 16 + 4 * g;
                cx<float> d[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    d[v] = mk(fr[u0 + v], fi[u0 + v]) * (row ? row[md.rec_d[l] + u0 + v] : mk(0.0f, 0.0f));
                    if (row) row[md.rec_d[l] + u0 + v] = d[v];
                }
                const int q = (mt & 1) * 2;
                dr[mt >> 1][q] = pack_bf16(d[0].re, d[1].re);
                dr[mt >> 1][q + 1] = pack_bf16(d[2].re, d[3].re);
                di[mt >> 1][q] = pack_bf16(d[0].im, d[1].im);
                di[mt >> 1][q + 1] = pack_bf16(d[2].im, d[3].im);
            }
        }
    }
    for (int l = md.L - 2; l >= 0; l--) {
        const int MT = md.out_p[l] >> 4, KB = (md.out_p[l + 1] + 31) >> 5;
        const uint4* Fl = F + md.bfrag_off[l];
        uint32_t nr[KBMAX][4], ni[KBMAX][4];
#pragma unroll
        for (int kb = 0; kb < KBMAX; kb++)
#pragma unroll
            for (int q = 0; q < 4; q++) { nr[kb][q] = 0u; ni[kb][q] = 0u; }
#pragma unroll
        for (int mt = 0; mt < 2 * KBMAX; mt++) {
            if (mt < MT) {
                const int u0 = mt * 16 + 4 * g;
                f32x4 ar = {0.0f, 0.0f, 0.0f, 0.0f}, ai = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kb = 0; kb < KBMAX; kb++) {
                    if (kb < KB) {
                        const uint4* fb = Fl + (size_t)((mt * KB + kb) * 2) * 64 + lane;
                        const bf16x8 Ar = as_bf16x8(fb[0]);
                        const bf16x8 Ai = as_bf16x8(fb[64]);
                        const bf16x8 Br = as_bf16x8(make_uint4(dr[kb][0], dr[kb][1], dr[kb][2], dr[kb][3]));
                        const bf16x8 Bi = as_bf16x8(make_uint4(di[kb][0], di[kb][1], di[kb][2], di[kb][3]));
                        const bf16x8 Bn = as_bf16x8(make_uint4(di[kb][0] ^ 0x80008000u, di[kb][1] ^ 0x80008000u,
                                                               di[kb][2] ^ 0x80008000u, di[kb][3] ^ 0x80008000u));
                        ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Br, ar, 0, 0, 0);
                        ar = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, Bn, ar, 0, 0, 0);
                        ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ai, Br, ai, 0, 0, 0);
                        ai = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ar, Bi, ai, 0, 0, 0);
                    }
                }
                cx<float> d[4];
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    d[v] = mk(ar[v], ai[v]) * (row ? row[md.rec_d[l] + u0 + v] : mk(0.0f, 0.0f));
                    if (row) row[md.rec_d[l] + u0 + v] = d[v];
                }
                const int q = (mt & 1) * 2;
                nr[mt >> 1][q] = pack_bf16(d[0].re, d[1].re);
                nr[mt >> 1][q + 1] = pack_bf16(d[2].re, d[3].re);
                ni[mt >> 1][q] = pack_bf16(d[0].im, d[1].im);
                ni[mt >> 1][q + 1] = pack_bf16(d[2].im, d[3].im);
            }
        }
#pragma unroll
        for (int kb = 0; kb < KBMAX; kb++)
#pragma unroll
            for (int q = 0; q < 4; q++) { dr[kb][q] = nr[kb][q]; di[kb][q] = ni[kb][q]; }
    }
}

// log psi and E_loc of given configurations (configs, or index_base + idx),
// 16 per wavefront, grid-stride over tiles.  Off-diagonal terms with a zero
// matrix element are skipped: each column walks its own list of non-zero
// terms and the wavefront runs as many forward passes as its longest list.
template <class FWD>
__global__ __launch_bounds__(512) void k_deep_mfma_eval(DeepMfmaDev md, OpDev op, int has_op, int64_t n,
                                                         const uint64_t* __restrict__ configs, uint64_t index_base,
                                                         Records rec) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const uint4* F;
    const float* V;
    mfma_stage(md, smem_raw, md.fwd_total, &F, &V);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 15;
    const int waves = blockDim.x >> 6;
    for (int64_t base = ((int64_t)blockIdx.x * waves + wave) * 16; base < n; base += (int64_t)gridDim.x * waves * 16) {
        const int64_t idx = base + col;
        const bool valid = idx < n;
        uint64_t s0 = 0, s1 = 0;
        if (valid) {
            if (configs) { s0 = configs[2 * idx]; s1 = configs[2 * idx + 1]; }
            else { s0 = index_base + (uint64_t)idx; s1 = 0; }
            mask_words(s0, s1, md.N);
        }
        const cx<float> lp = FWD::run(md, F, V, s0, s1, lane);
        cx<double> E = mk(0.0, 0.0);
        if (has_op) {
            E = diag_energy(op, s0, s1);
            int t = valid ? 0 : op.n_terms;
            for (;;) {
                cx<double> cf = mk(0.0, 0.0);
                bool act = false;
                while (t < op.n_terms) {
                    cf = term_coef_lane(op, t, s0, s1);
                    if (cf.re != 0.0 || cf.im != 0.0) { act = true; break; }
                    t++;
                }
                if (__ballot(act) == 0) break;
                uint64_t p0 = s0, p1 = s1;
                if (act)
                    for (int q = op.flip_begin[t]; q < op.flip_begin[t + 1]; q++) flip_bit(p0, p1, op.flips[q]);
                const cx<float> lpp = FWD::run(md, F, V, p0, p1, lane);
                if (act) {
                    E = E + cf * cexp_wide(lpp - lp);
                    t++;
                }
            }
        }
        if (valid && lane < 16) {
            if (rec.spins) { rec.spins[2 * idx] = s0; rec.spins[2 * idx + 1] = s1; }
            if (rec.log_psi) rec.log_psi[idx] = to_d(lp);
            if (rec.e_loc) rec.e_loc[idx] = E;
        }
    }
}

// Metropolis single-flip chains, 16 per wavefront, with the full forward pass
// recomputed per proposal (MonteCarloLoop.hpp:153-256 with PsiDeep.hpp:191-194)
// and the proposal stream of the other PsiDeep sampler: Philox block
// (chain, k/2, call, 0) gives site and uniform of proposals k and k+1.  Only
// the configurations are recorded; the evaluation kernel measures them.
template <class FWD>
__global__ __launch_bounds__(512) void k_deep_mfma_mc(DeepMfmaDev md, McParams mp, uint64_t* __restrict__ out_spins,
                                                       uint32_t* __restrict__ counts) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const uint4* F;
    const float* V;
    mfma_stage(md, smem_raw, md.fwd_total, &F, &V);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 15;
    const int waves = blockDim.x >> 6;
    const int chain = (blockIdx.x * waves + wave) * 16 + col;
    const bool valid = chain < mp.chains;
    const int N = md.N;
    uint64_t s0 = 0, s1 = 0;
    if (valid) {
        const u32x4 r0 = philox((uint32_t)chain, 0u, mp.call, 1u, mp.seed_lo, mp.seed_hi);
        s0 = (uint64_t)r0.x | ((uint64_t)r0.y << 32);
        s1 = (uint64_t)r0.z | ((uint64_t)r0.w << 32);
        mask_words(s0, s1, N);
        if (mp.move == MC_EXCHANGE) sector_start(N, mp.k_up, (uint32_t)chain, mp.call, mp.seed_lo, mp.seed_hi, s0, s1);
    }
    float cur = FWD::run(md, F, V, s0, s1, lane).re;
    uint32_t nacc = 0, nrej = 0;
    const uint64_t n_therm = (uint64_t)mp.therm * N;
    const uint64_t n_sweep = (uint64_t)mp.sweeps * (mp.move == MC_FAST_SWEEP ? 1 : N);
    const uint64_t total = n_therm + n_sweep * (uint64_t)mp.steps_per_chain;
    uint64_t next_rec = n_therm + n_sweep;
    int step = 0;
    // records that fall before the first proposal (no thermalisation, zero sweeps)
    while (step < mp.steps_per_chain && next_rec == 0) {
        if (valid && lane < 16) {
            const int64_t sample = (int64_t)step * mp.chains + chain;
            out_spins[2 * sample] = s0;
            out_spins[2 * sample + 1] = s1;
        }
        step++;
        next_rec += n_sweep;
    }
    uint32_t rx = 0, ry = 0, rz = 0, rw = 0;
    for (uint64_t k = 0; k < total; k++) {
        uint64_t n0 = s0, n1 = s1;
        uint32_t y;
        if (mp.move == MC_SINGLE_FLIP) {
            if ((k & 1) == 0) {
                const u32x4 r = philox((uint32_t)chain, (uint32_t)(k >> 1), mp.call, 0u, mp.seed_lo, mp.seed_hi);
                rx = r.x; ry = r.y; rz = r.z; rw = r.w;
            }
            const uint32_t x = (k & 1) ? rz : rx;
            y = (k & 1) ? rw : ry;
            flip_bit(n0, n1, (int)(x % (uint32_t)N));
        } else if (valid) {
            propose_move(mp.move, mp.tries, N, (uint32_t)chain, k, mp.call, mp.seed_lo, mp.seed_hi, s0, s1, n0, n1, y);
        } else {
            y = 0u;
        }
        const float next = FWD::run(md, F, V, n0, n1, lane).re;
        const float ratio = expf(2.0f * (next - cur));
        const float uu = (float)(((double)y + 0.5) * (1.0 / 4294967296.0));
        if ((ratio > 1.0f) || (uu <= ratio)) { s0 = n0; s1 = n1; cur = next; nacc++; }
        else nrej++;
        while (step < mp.steps_per_chain && k + 1 == next_rec) {
            if (valid && lane < 16) {
                const int64_t sample = (int64_t)step * mp.chains + chain;
                out_spins[2 * sample] = s0;
                out_spins[2 * sample + 1] = s1;
            }
            step++;
            next_rec += n_sweep;
        }
    }
    if (valid && lane < 16 && counts) { counts[2 * chain] = nacc; counts[2 * chain + 1] = nrej; }
}

// Gradient records of given configurations: the forward pass of log psi with
// x_l and my_tanh(a_l) written to the configuration's row, then the backward
// pass (16 configurations per wavefront, grid-stride over tiles).
template <int KBMAX>
__global__ __launch_bounds__(512) void k_deep_mfma_grad(DeepMfmaDev md, int64_t n, const uint64_t* __restrict__ configs,
                                                        cx<float>* __restrict__ records) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const uint4* F;
    const float* V;
    mfma_stage(md, smem_raw, md.frag_total, &F, &V);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 15;
    const int waves = blockDim.x >> 6;
    for (int64_t base = ((int64_t)blockIdx.x * waves + wave) * 16; base < n; base += (int64_t)gridDim.x * waves * 16) {
        const int64_t idx = base + col;
        const bool valid = idx < n;
        uint64_t s0 = 0, s1 = 0;
        cx<float>* row = nullptr;
        if (valid) {
            s0 = configs[2 * idx];
            s1 = configs[2 * idx + 1];
            mask_words(s0, s1, md.N);
            row = records + idx * md.rec_stride;
        }
        const cx<float> lp = mfma_log_psi<KBMAX, true>(md, F, V, s0, s1, lane, row);
        mfma_backward<KBMAX>(md, F, V, row, lane);
        if (valid && lane < 16) row[md.rec_lp] = lp;
    }
}

// Rows of the contraction GEMM G = X^T Y for one layer (see deep_kernels.hpp).
// A workgroup writes whole sample rows (grid-stride over the rows), so the
// row's coefficient, spin words and record pointer are read once per row and
// the columns need no division.
__global__ __launch_bounds__(256) void k_deep_grad_rows(DeepMfmaDev md, int layer, int in, int out, int64_t s0, int64_t cnt,
                                                        int64_t rows, const uint64_t* __restrict__ configs,
                                                        const cx<float>* __restrict__ records,
                                                        const cx<double>* __restrict__ coef, int lp_factor,
                                                        float* __restrict__ X, int ldx, float* __restrict__ Y, int ldy) {
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        float* xr = X + r * ldx;
        float* yr = Y + r * ldy;
        if (r >= cnt) {   // padding rows of the last K block
            for (int c = threadIdx.x; c < ldx; c += blockDim.x) xr[c] = 0.0f;
            for (int c = threadIdx.x; c < ldy; c += blockDim.x) yr[c] = 0.0f;
            continue;
        }
        const int64_t s = s0 + r;
        const cx<float>* row = records + s * md.rec_stride;
        if (layer == 0) {
            const uint64_t w0 = configs[2 * s], w1 = configs[2 * s + 1];
            for (int c = threadIdx.x; c < ldx; c += blockDim.x)
                xr[c] = c < in ? spin_of<float>(w0, w1, c) : (c == 2 * in ? 1.0f : 0.0f);
        } else {
            const cx<float>* x = row + md.rec_x[layer - 1];
            for (int c = threadIdx.x; c < ldx; c += blockDim.x)
                xr[c] = c < in ? x[c].re : (c < 2 * in ? x[c - in].im : (c == 2 * in && layer < md.L ? 1.0f : 0.0f));
        }
        const cx<double> cf = conj(coef[s]);
        if (layer < md.L) {
            const cx<float>* d = row + md.rec_d[layer];
            for (int c = threadIdx.x; c < ldy; c += blockDim.x) {
                float v = 0.0f;
                if (c < 2 * out) {
                    const cx<double> y = cf * to_d(d[c < out ? c : c - out]);
                    v = (float)(c < out ? y.re : y.im);
                }
                yr[c] = v;
            }
        } else {
            const cx<double> y = lp_factor ? cf * to_d(row[md.rec_lp]) : cf;
            for (int c = threadIdx.x; c < ldy; c += blockDim.x) yr[c] = c == 0 ? (float)y.re : (c == 1 ? (float)y.im : 0.0f);
        }
    }
}

size_t deep_mfma_lds_bytes(const DeepMfmaDev& md, bool grad) {
    return (size_t)(grad ? md.frag_total : md.fwd_total) * 16 + (size_t)md.vec_total * 4;
}

int deep_mfma_kbmax(const DeepMfmaDev& md) {
    int kb = 0;
    for (int l = 0; l < md.L; l++) {
        kb = md.in_p[l] / 32 > kb ? md.in_p[l] / 32 : kb;
        kb = (md.out_p[l] + 31) / 32 > kb ? (md.out_p[l] + 31) / 32 : kb;
    }
    return kb;
}

// K blocks of the widest layer, rounded up to an instantiated register budget
#define RBM_DM_DISPATCH(kbm, X) \
    if ((kbm) <= 2) { X(2) }    \
    else if ((kbm) <= 4) { X(4) } \
    else if ((kbm) <= 5) { X(5) } \
    else if ((kbm) <= 6) { X(6) } \
    else { X(8) }

template <typename K>
static hipError_t set_lds(K kernel, size_t lds) {
    if (lds > 65536) return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return hipSuccess;
}

hipError_t launch_deep_mfma_eval(const DeepMfmaDev& md, const OpDev& op, int has_op, int64_t n, const uint64_t* configs,
                                 uint64_t index_base, const Records& rec, int num_cus, hipStream_t stream) {
    const size_t lds = deep_mfma_lds_bytes(md);
    const int kbm = deep_mfma_kbmax(md);
    const int64_t tiles = (n + 15) / 16;
    const int waves = 8;
    int64_t blocks = (tiles + waves - 1) / waves;
    if (blocks > num_cus) blocks = num_cus;
    if (blocks < 1) blocks = 1;
    hipError_t e = hipSuccess;
#define RBM_DM_EVAL_F(FW)                                                                                          \
    if ((e = set_lds(&k_deep_mfma_eval<FW>, lds)) != hipSuccess) return e;                                        \
    hipLaunchKernelGGL((k_deep_mfma_eval<FW>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, md, op, has_op, n, \
                       configs, index_base, rec);
#define RBM_DM_EVAL(KB) RBM_DM_EVAL_F(FwdAny<KB>)
    if (is_shape_40x160x80(md)) {
        RBM_DM_EVAL_F(FwdDense40x160x80)
    } else {
        RBM_DM_DISPATCH(kbm, RBM_DM_EVAL)
    }
#undef RBM_DM_EVAL
#undef RBM_DM_EVAL_F
    return hipGetLastError();
}

hipError_t launch_deep_mfma_mc(const DeepMfmaDev& md, const McParams& mp, uint64_t* spins, uint32_t* counts, int num_cus,
                               hipStream_t stream) {
    const size_t lds = deep_mfma_lds_bytes(md);
    const int kbm = deep_mfma_kbmax(md);
    const int wave_count = (mp.chains + 15) / 16;
    // one workgroup per CU holds the weights; spread the waves over all CUs
    int waves = (wave_count + num_cus - 1) / num_cus;
    if (waves < 1) waves = 1;
    if (waves > 8) waves = 8;
    const int blocks = (wave_count + waves - 1) / waves;
    hipError_t e = hipSuccess;
#define RBM_DM_MC_F(FW)                                                                                            \
    if ((e = set_lds(&k_deep_mfma_mc<FW>, lds)) != hipSuccess) return e;                                          \
    hipLaunchKernelGGL((k_deep_mfma_mc<FW>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, md, mp, spins, counts);
#define RBM_DM_MC(KB) RBM_DM_MC_F(FwdAny<KB>)
    if (is_shape_40x160x80(md)) {
        RBM_DM_MC_F(FwdDense40x160x80)
    } else {
        RBM_DM_DISPATCH(kbm, RBM_DM_MC)
    }
#undef RBM_DM_MC
#undef RBM_DM_MC_F
    return hipGetLastError();
}

hipError_t launch_deep_mfma_grad(const DeepMfmaDev& md, int64_t n, const uint64_t* configs, cx<float>* records,
                                 int num_cus, hipStream_t stream) {
    const size_t lds = deep_mfma_lds_bytes(md, true);
    const int kbm = deep_mfma_kbmax(md);
    const int64_t tiles = (n + 15) / 16;
    const int waves = 8;
    int64_t blocks = (tiles + waves - 1) / waves;
    if (blocks > num_cus) blocks = num_cus;
    if (blocks < 1) blocks = 1;
    hipError_t e = hipSuccess;
#define RBM_DM_GRAD(KB)                                                                                            \
    if ((e = set_lds(&k_deep_mfma_grad<KB>, lds)) != hipSuccess) return e;                                        \
    hipLaunchKernelGGL((k_deep_mfma_grad<KB>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, md, n, configs, records);
    RBM_DM_DISPATCH(kbm, RBM_DM_GRAD)
#undef RBM_DM_GRAD
    return hipGetLastError();
}

hipError_t launch_deep_grad_rows(const DeepMfmaDev& md, int layer, int in, int out, int64_t s0, int64_t cnt, int64_t rows,
                                 const uint64_t* configs, const cx<float>* records, const cx<double>* coef,
                                 int lp_factor, float* X, int ldx, float* Y, int ldy, hipStream_t stream) {
    int64_t blocks = rows < 65536 ? rows : 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_deep_grad_rows, dim3((unsigned)blocks), dim3(256), 0, stream, md, layer, in, out, s0, cnt, rows,
                       configs, records, coef, lp_factor, X, ldx, Y, ldy);
    return hipGetLastError();
}

}  // namespace rbmhip
