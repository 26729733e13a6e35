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
template <int KBMAX>
RBM_DI cx<float> mfma_log_psi(const DeepMfmaDev& md, const uint4* __restrict__ F, const float* __restrict__ V,
                              uint64_t s0, uint64_t s1, int lane) {
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

// Copy the fragments and vectors into LDS (dynamic shared memory).
RBM_DI void mfma_stage(const DeepMfmaDev& md, unsigned char* smem, const uint4** F, const float** V) {
    uint4* f = reinterpret_cast<uint4*>(smem);
    for (int i = threadIdx.x; i < md.frag_total; i += blockDim.x) f[i] = md.frag[i];
    float* v = reinterpret_cast<float*>(f + md.frag_total);
    for (int i = threadIdx.x; i < md.vec_total; i += blockDim.x) v[i] = md.vec[i];
    __syncthreads();
    *F = f;
    *V = v;
}

// log psi and E_loc of given configurations (configs, or index_base + idx),
// 16 per wavefront, grid-stride over tiles.  Off-diagonal terms with a zero
// matrix element are skipped: each column walks its own list of non-zero
// terms and the wavefront runs as many forward passes as its longest list.
template <int KBMAX>
__global__ __launch_bounds__(512) void k_deep_mfma_eval(DeepMfmaDev md, OpDev op, int has_op, int64_t n,
                                                         const uint64_t* __restrict__ configs, uint64_t index_base,
                                                         Records rec) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const uint4* F;
    const float* V;
    mfma_stage(md, smem_raw, &F, &V);
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
        const cx<float> lp = mfma_log_psi<KBMAX>(md, F, V, s0, s1, lane);
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
                const cx<float> lpp = mfma_log_psi<KBMAX>(md, F, V, p0, p1, lane);
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
template <int KBMAX>
__global__ __launch_bounds__(512) void k_deep_mfma_mc(DeepMfmaDev md, McParams mp, uint64_t* __restrict__ out_spins,
                                                       uint32_t* __restrict__ counts) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const uint4* F;
    const float* V;
    mfma_stage(md, smem_raw, &F, &V);
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
    float cur = mfma_log_psi<KBMAX>(md, F, V, s0, s1, lane).re;
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
        const float next = mfma_log_psi<KBMAX>(md, F, V, n0, n1, lane).re;
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

size_t deep_mfma_lds_bytes(const DeepMfmaDev& md) {
    return (size_t)md.frag_total * 16 + (size_t)md.vec_total * 4;
}

int deep_mfma_kbmax(const DeepMfmaDev& md) {
    int kb = 0;
    for (int l = 0; l < md.L; l++) {
        kb = md.in_p[l] / 32 > kb ? md.in_p[l] / 32 : kb;
        kb = (md.out_p[l] + 31) / 32 > kb ? (md.out_p[l] + 31) / 32 : kb;
    }
    return kb;
}

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
    hipError_t e;
    if (kbm <= 4) {
        if ((e = set_lds(&k_deep_mfma_eval<4>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_deep_mfma_eval<4>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, md, op, has_op, n,
                           configs, index_base, rec);
    } else {
        if ((e = set_lds(&k_deep_mfma_eval<8>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_deep_mfma_eval<8>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, md, op, has_op, n,
                           configs, index_base, rec);
    }
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
    hipError_t e;
    if (kbm <= 4) {
        if ((e = set_lds(&k_deep_mfma_mc<4>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_deep_mfma_mc<4>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, md, mp, spins, counts);
    } else {
        if ((e = set_lds(&k_deep_mfma_mc<8>, lds)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_deep_mfma_mc<8>), dim3((unsigned)blocks), dim3(64 * waves), lds, stream, md, mp, spins, counts);
    }
    return hipGetLastError();
}

}  // namespace rbmhip
