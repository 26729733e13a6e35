// gfx950 kernels for PsiDeep (include/quantum_state/PsiDeep.hpp): sparse
// complex layers with my_logcosh activations and a weighted output sum.
//
// A 16-lane row owns one configuration (one Markov chain); unit j of a layer
// is computed by lane j % 16.  The row keeps its activations, pre-activations
// and back-propagated derivatives in its own slice of LDS; a workgroup is one
// wavefront (four rows), so __syncthreads() is a wave-local fence.
#include "deep_kernels.hpp"

namespace rbmhip {

template <typename T>
struct RowLds {
    cx<T>* a0;
    cx<T>* a1;
    cx<T>* ang;
    cx<T>* d0;
    cx<T>* d1;
};

template <typename T>
RBM_HDI size_t deep_lds_per_row(int width, int units) {
    return (size_t)(4 * width + units) * sizeof(cx<T>);
}

template <typename T>
RBM_DI RowLds<T> row_lds(unsigned char* smem, int row, int width, int units) {
    cx<T>* base = reinterpret_cast<cx<T>*>(smem) + (size_t)row * (4 * width + units);
    RowLds<T> g;
    g.a0 = base;
    g.a1 = base + width;
    g.ang = base + 2 * width;
    g.d0 = base + 2 * width + units;
    g.d1 = base + 3 * width + units;
    return g;
}

// my_logcosh(a) as a value (not a difference): the cancellation-free form
// plus its constant, so that a small activation (~a^2 / 2) keeps its relative
// accuracy.  The gradient records (activations as inputs of the next layer's
// weight derivatives, and the final-weight derivatives) use it; log psi
// differences of E_loc and the sampler keep the direct form, whose rounding
// is common to both sides of a difference.
template <typename T>
RBM_DI cx<T> logcosh_value(cx<T> a) {
    const cx<T> s = logcosh_shifted(a);
    return mk(s.re + T(LC_L0), s.im);
}

// forward_pass (PsiDeep.hpp:89-133) on the configuration rotated left by
// `shift` (Spins.h:280-285: new bit k = old bit (k - shift) mod N).  Returns
// this lane's share of sum_j x_j final_j; stores pre-activations if asked
// (the gradient records' forward pass, which then evaluates the activations
// with logcosh_value).
template <typename T>
RBM_DI cx<T> deep_forward(const DeepDev<T>& md, uint64_t s0, uint64_t s1, int shift, const RowLds<T>& g,
                          bool store_ang, int lane) {
    const int N = md.N;
    __syncthreads();
    for (int i = lane; i < N; i += 16) {
        int src = i - shift;
        if (src < 0) src += N;
        g.a0[i] = mk(spin_of<T>(s0, s1, src), T(0));
    }
    __syncthreads();
    cx<T>* x = g.a0;
    cx<T>* y = g.a1;
    int abegin = 0;
    for (int l = 0; l < md.num_layers; l++) {
        const int size = md.size[l], C = md.conn[l];
        const uint32_t* cn = md.conns + md.w_off[l];
        const cx<T>* w = md.weights + md.w_off[l];
        const cx<T>* b = md.biases + md.b_off[l];
        for (int j = lane; j < size; j += 16) {
            cx<T> a = mk(T(0), T(0));
            for (int i = 0; i < C; i++) a = a + w[i * size + j] * x[cn[i * size + j]];
            a = a + b[j];
            if (store_ang) g.ang[abegin + j] = a;
            y[j] = store_ang ? logcosh_value(a) : logcosh(a);
        }
        __syncthreads();
        cx<T>* t = x; x = y; y = t;
        abegin += size;
    }
    const int last = md.size[md.num_layers - 1];
    cx<T> part = mk(T(0), T(0));
    for (int j = lane; j < last; j += 16) part = part + x[j] * md.final_w[j];
    return part;
}

// log_psi_s_generic (PsiDeep.hpp:135-169): mean over the N cyclic shifts when
// translational invariance is on.
template <typename T>
RBM_DI cx<T> deep_log_psi(const DeepDev<T>& md, uint64_t s0, uint64_t s1, const RowLds<T>& g, int lane) {
    const int shifts = md.translational_invariance ? md.N : 1;
    cx<T> acc = mk(T(0), T(0));
    for (int sh = 0; sh < shifts; sh++) acc = acc + deep_forward(md, s0, s1, sh, g, false, lane);
    cx<T> r = row_sum16(acc);
    if (md.translational_invariance) r = (T(1) / T(md.N)) * r;
    return r;
}

template <typename T>
RBM_DI cx<double> local_energy_deep(const DeepDev<T>& md, const OpDev& op, uint64_t s0, uint64_t s1, cx<T> lp,
                                    const RowLds<T>& g, int lane) {
    cx<double> E = diag_energy(op, s0, s1);
    for (int t = 0; t < op.n_terms; t++) {
        uint64_t p0 = s0, p1 = s1;
        for (int q = op.flip_begin[t]; q < op.flip_begin[t + 1]; q++) flip_bit(p0, p1, op.flips[q]);
        const cx<T> lpp = deep_log_psi(md, p0, p1, g, lane);
        E = E + term_coef(op, t, s0, s1) * cexp_wide(lpp - lp);
    }
    return E;
}

// foreach_O_k (PsiDeep.hpp:214-293): back-propagation through the layers on
// the un-shifted configuration.  Writes the K = num_params - 2N derivatives in
// parameter order [(b_l, W_l) per layer, final weights].
template <typename T>
RBM_DI void deep_o_k(const DeepDev<T>& md, uint64_t s0, uint64_t s1, const RowLds<T>& g, int lane, cx<T>* Orow) {
    const cx<T> lp = row_sum16(deep_forward(md, s0, s1, 0, g, true, lane));
    const int L = md.num_layers;
    int64_t bp[DEEP_MAX_LAYERS];
    int ab[DEEP_MAX_LAYERS];
    int64_t p = 0;
    int a = 0;
    for (int l = 0; l < L; l++) {
        bp[l] = p;
        ab[l] = a;
        p += md.size[l] + (int64_t)md.conn[l] * md.size[l];
        a += md.size[l];
    }
    cx<T>* D = g.d0;
    cx<T>* ND = g.d1;
    __syncthreads();
    for (int l = L - 1; l >= 0; l--) {
        const int size = md.size[l], C = md.conn[l];
        if (l == L - 1) {
            for (int j = lane; j < size; j += 16) D[j] = md.final_w[j] * dlogcosh(g.ang[ab[l] + j]);
        } else {
            const int nsize = md.size[l + 1];
            const cx<T>* nw = md.weights + md.w_off[l + 1];
            const int* rb = md.rhs_begin + md.rb_off[l];
            for (int i = lane; i < size; i += 16) {
                cx<T> acc = mk(T(0), T(0));
                for (int e = rb[i]; e < rb[i + 1]; e++) {
                    const int slot = md.rhs_entries[e];
                    acc = acc + nw[slot] * D[slot % nsize];
                }
                ND[i] = acc * dlogcosh(g.ang[ab[l] + i]);
            }
            __syncthreads();
            cx<T>* t = D; D = ND; ND = t;
        }
        __syncthreads();
        const uint32_t* cn = md.conns + md.w_off[l];
        for (int j = lane; j < size; j += 16) {
            Orow[bp[l] + j] = D[j];
            for (int c = 0; c < C; c++) {
                const uint32_t u = cn[c * size + j];
                const cx<T> xp = l == 0 ? mk(spin_of<T>(s0, s1, (int)u), T(0)) : logcosh_value(g.ang[ab[l - 1] + u]);
                Orow[bp[l] + size + (int64_t)c * size + j] = D[j] * xp;
            }
        }
    }
    const int last = md.size[L - 1];
    for (int j = lane; j < last; j += 16) {
        const cx<T> x = logcosh_value(g.ang[ab[L - 1] + j]);
        Orow[p + j] = md.ref_compat_final ? lp * x : x;
    }
    __syncthreads();
}

template <typename T>
RBM_DI void deep_measure(const DeepDev<T>& md, const OpDev& op, int has_op, uint64_t s0, uint64_t s1,
                         const RowLds<T>& g, int lane, int64_t sample, const Records& rec, int64_t K) {
    const cx<T> lp = deep_log_psi(md, s0, s1, g, lane);
    cx<double> E = mk(0.0, 0.0);
    if (has_op) E = local_energy_deep(md, op, s0, s1, lp, g, lane);
    if (rec.tanh) deep_o_k(md, s0, s1, g, lane, reinterpret_cast<cx<T>*>(rec.tanh) + sample * K);
    if (lane == 0) {
        if (rec.spins) { rec.spins[2 * sample] = s0; rec.spins[2 * sample + 1] = s1; }
        if (rec.log_psi) rec.log_psi[sample] = to_d(lp);
        if (rec.e_loc) rec.e_loc[sample] = E;
    }
}

template <typename T>
RBM_DI int64_t deep_K(const DeepDev<T>& md) {
    int64_t p = 0;
    for (int l = 0; l < md.num_layers; l++) p += md.size[l] + (int64_t)md.conn[l] * md.size[l];
    return p + md.size[md.num_layers - 1];
}

template <typename T>
__global__ __launch_bounds__(64) void k_deep_eval(DeepDev<T> md, OpDev op, int has_op, int64_t n,
                                                  const uint64_t* __restrict__ configs, uint64_t index_base,
                                                  Records rec) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = threadIdx.x & 15, row = threadIdx.x >> 4;
    const RowLds<T> g = row_lds<T>(smem_raw, row, md.width, md.units);
    const int64_t K = deep_K(md);
    for (int64_t idx = (int64_t)blockIdx.x * 4 + row; idx < n; idx += (int64_t)gridDim.x * 4) {
        uint64_t s0, s1;
        if (configs) { s0 = configs[2 * idx]; s1 = configs[2 * idx + 1]; }
        else { s0 = index_base + (uint64_t)idx; s1 = 0; }
        mask_words(s0, s1, md.N);
        deep_measure(md, op, has_op, s0, s1, g, lane, idx, rec, K);
    }
}

// Metropolis chains for PsiDeep with the same proposal stream as the RBM
// sampler; the full forward pass is recomputed per proposal
// (PsiDeep.hpp:191-194 makes the angle update a no-op).
template <typename T>
__global__ __launch_bounds__(64) void k_deep_mc(DeepDev<T> md, OpDev op, int has_op, McParams mp, Records rec,
                                                uint32_t* __restrict__ counts) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int lane = threadIdx.x & 15, row = threadIdx.x >> 4;
    const int chain = blockIdx.x * 4 + row;
    if (chain >= mp.chains) return;
    const RowLds<T> g = row_lds<T>(smem_raw, row, md.width, md.units);
    const int N = md.N;
    const int64_t K = deep_K(md);
    const u32x4 r0 = philox((uint32_t)chain, 0u, mp.call, 1u, mp.seed_lo, mp.seed_hi);
    uint64_t s0 = (uint64_t)r0.x | ((uint64_t)r0.y << 32);
    uint64_t s1 = (uint64_t)r0.z | ((uint64_t)r0.w << 32);
    mask_words(s0, s1, N);
    if (mp.move == MC_EXCHANGE) sector_start(N, mp.k_up, (uint32_t)chain, mp.call, mp.seed_lo, mp.seed_hi, s0, s1);
    T cur = deep_log_psi(md, s0, s1, g, lane).re;
    uint32_t nacc = 0, nrej = 0;
    u32x4 rr = r0;
    uint64_t k = 0;
    auto propose = [&](uint64_t kk) {
        uint64_t n0 = s0, n1 = s1;
        uint32_t y;
        if (mp.move == MC_SINGLE_FLIP) {
            if ((kk & 1) == 0) rr = philox((uint32_t)chain, (uint32_t)(kk >> 1), mp.call, 0u, mp.seed_lo, mp.seed_hi);
            const uint32_t x = (kk & 1) ? rr.z : rr.x;
            y = (kk & 1) ? rr.w : rr.y;
            flip_bit(n0, n1, (int)(x % (uint32_t)N));
        } else {
            propose_move(mp.move, mp.tries, N, (uint32_t)chain, kk, mp.call, mp.seed_lo, mp.seed_hi, s0, s1, n0, n1, y);
        }
        const T next = deep_log_psi(md, n0, n1, g, lane).re;
        const T ratio = (T)exp(T(2) * (next - cur));
        const T uu = (T)(((double)y + 0.5) * (1.0 / 4294967296.0));
        if ((ratio > T(1)) || (uu <= ratio)) { s0 = n0; s1 = n1; cur = next; nacc++; }
        else nrej++;
    };
    const uint64_t n_therm = (uint64_t)mp.therm * N;
    const uint64_t n_sweep = (uint64_t)mp.sweeps * (mp.move == MC_FAST_SWEEP ? 1 : N);
    for (uint64_t i = 0; i < n_therm; i++) propose(k++);
    for (int step = 0; step < mp.steps_per_chain; step++) {
        for (uint64_t i = 0; i < n_sweep; i++) propose(k++);
        deep_measure(md, op, has_op, s0, s1, g, lane, (int64_t)step * mp.chains + chain, rec, K);
    }
    if (lane == 0 && counts) { counts[2 * chain] = nacc; counts[2 * chain + 1] = nrej; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_rows_reduce(int64_t n, int64_t K, const cx<T>* __restrict__ O,
                                                     const cx<double>* __restrict__ coef, int nc, int64_t chunk,
                                                     cx<double>* __restrict__ partials) {
    const int64_t s_begin = (int64_t)blockIdx.x * chunk;
    const int64_t s_end = s_begin + chunk < n ? s_begin + chunk : n;
    for (int64_t k = threadIdx.x; k < K; k += blockDim.x) {
        cx<double> a0 = mk(0.0, 0.0), a1 = mk(0.0, 0.0);
        for (int64_t s = s_begin; s < s_end; s++) {
            const cx<double> o = conj(to_d(O[s * K + k]));
            a0 = a0 + coef[s] * o;
            if (nc > 1) a1 = a1 + coef[n + s] * o;
        }
        partials[(int64_t)blockIdx.x * nc * K + k] = a0;
        if (nc > 1) partials[(int64_t)blockIdx.x * nc * K + K + k] = a1;
    }
}

__global__ void k_exp_vector(int64_t n, const cx<double>* __restrict__ lp, double lpf, cx<double>* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = cexp(lp[i] + lpf);
}

template <typename T>
hipError_t launch_deep_eval(const DeepDev<T>& md, const OpDev& op, int has_op, int64_t n, const uint64_t* configs,
                            uint64_t index_base, const Records& rec, int max_blocks, hipStream_t stream) {
    int64_t blocks = (n + 3) / 4;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    const size_t shm = 4 * deep_lds_per_row<T>(md.width, md.units);
    if (shm > 65536) (void)hipFuncSetAttribute((const void*)&k_deep_eval<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL((k_deep_eval<T>), dim3((unsigned)blocks), dim3(64), shm, stream, md, op, has_op, n, configs,
                       index_base, rec);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_deep_mc(const DeepDev<T>& md, const OpDev& op, int has_op, const McParams& mp, const Records& rec,
                          uint32_t* counts, hipStream_t stream) {
    const size_t shm = 4 * deep_lds_per_row<T>(md.width, md.units);
    if (shm > 65536) (void)hipFuncSetAttribute((const void*)&k_deep_mc<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL((k_deep_mc<T>), dim3((mp.chains + 3) / 4), dim3(64), shm, stream, md, op, has_op, mp, rec, counts);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_rows_reduce(int64_t n, int64_t K, const void* O, const cx<double>* coef, int nc, int64_t chunk,
                              int nblocks, cx<double>* partials, hipStream_t stream) {
    hipLaunchKernelGGL((k_rows_reduce<T>), dim3(nblocks), dim3(256), 0, stream, n, K,
                       reinterpret_cast<const cx<T>*>(O), coef, nc, chunk, partials);
    return hipGetLastError();
}

hipError_t launch_exp_vector(int64_t n, const cx<double>* lp, double log_prefactor, cx<double>* out, hipStream_t stream) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_exp_vector, dim3((unsigned)g), dim3(256), 0, stream, n, lp, log_prefactor, out);
    return hipGetLastError();
}

template hipError_t launch_deep_eval<float>(const DeepDev<float>&, const OpDev&, int, int64_t, const uint64_t*, uint64_t, const Records&, int, hipStream_t);
template hipError_t launch_deep_eval<double>(const DeepDev<double>&, const OpDev&, int, int64_t, const uint64_t*, uint64_t, const Records&, int, hipStream_t);
template hipError_t launch_deep_mc<float>(const DeepDev<float>&, const OpDev&, int, const McParams&, const Records&, uint32_t*, hipStream_t);
template hipError_t launch_deep_mc<double>(const DeepDev<double>&, const OpDev&, int, const McParams&, const Records&, uint32_t*, hipStream_t);
template hipError_t launch_rows_reduce<float>(int64_t, int64_t, const void*, const cx<double>*, int, int64_t, int, cx<double>*, hipStream_t);
template hipError_t launch_rows_reduce<double>(int64_t, int64_t, const void*, const cx<double>*, int, int64_t, int, cx<double>*, hipStream_t);

}  // namespace rbmhip
