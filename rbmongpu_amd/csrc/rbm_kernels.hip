// gfx950 kernels for the RBM wavefunction (reference Psi):
//   k_mc_rbm    Metropolis single-spin-flip chains with fused measurement
//   k_eval_rbm  log psi / E_loc / tanh(theta) for given configurations
//               (ExactSummation, psi_vector, psi_O_k_vector, batched calls)
//   k_ok_reduce + k_partials_sum   FP64 O_k contractions
//   k_coef, k_scalar_partials, k_lse_partials   per-sample FP64 reductions
//
// Execution model: a chain (or a basis state) is owned by a 16-lane DPP row;
// a wavefront holds four chains and a 256-thread workgroup sixteen.  Hidden
// unit j lives on lane j % 16 in slot j / 16, so theta_j stays in VGPRs and a
// row of W (contiguous in LDS) is read conflict-free.  Sums over hidden units
// are row rotations (DPP), never LDS round trips or __syncthreads.
#include "rbm_device.hpp"
#include "rbm_kernels.hpp"

namespace rbmhip {

// -2 s_p from a 32-bit word shifted so that bit p is its sign bit; the FP32
// value is assembled from bits (+2 or -2).
RBM_DI float flip_factor(uint32_t top) { return __uint_as_float((top & 0x80000000u) | 0x40000000u); }
template <typename T> RBM_DI T flip_factor_t(uint32_t top) { return (top & 0x80000000u) ? T(-2) : T(2); }
template <> RBM_DI float flip_factor_t<float>(uint32_t top) { return flip_factor(top); }

// -2 s_p for a wave-uniform site p (an operator table entry): the word is
// chosen by uniform tests, with no 64-bit shifts.
template <typename T> RBM_DI T spin_factor(uint64_t s0, uint64_t s1, int p) {
    const uint64_t w64 = p < 64 ? s0 : s1;
    const uint32_t w = (p & 32) ? (uint32_t)(w64 >> 32) : (uint32_t)w64;
    return flip_factor_t<T>(w << (31 - (p & 31)));
}

// ---------------------------------------------------------------------------
// local energy (include/operator/Operator.hpp:90-115, restated on canonical
// flip-mask terms): E_loc(s) = diag(s) + sum_t c_t(s) exp(log psi(s^m_t) - log psi(s)).
// Each lane sums lc(theta'_j) over its hidden units and subtracts lsum, the
// same sum at theta (padded units, theta = theta' = 0, cancel); the row adds
// the lanes.  Every off-diagonal term flips at least one spin, so theta' is
// initialised from the first flip without a copy of theta.  The per-term
// tail (coefficient, exp of the log ratio, FP64 accumulation) is spread over
// the row: lane t mod 16 keeps term t's log ratio, and after every 16 terms
// each lane finishes its own term; the diagonal strings are dealt out the same
// way, and the lanes' sums are added at the end.
// ---------------------------------------------------------------------------
template <typename T, int JPL>
RBM_DI cx<double> local_energy_rbm(const RbmDev<T>& md, const cx<T>* __restrict__ Wsrc, const OpDev& op,
                                   uint64_t s0, uint64_t s1, const cx<T> (&th)[JPL], cx<T> lsum, int lane) {
    const int ld = md.ld;
    const cx<T>* __restrict__ Wl = Wsrc + lane;
    cx<double> E = mk(0.0, 0.0);          // this lane's terms
    // diagonal strings d = lane, lane + 16, ... of this lane
    for (int d = lane; d < op.n_diag; d += 16)
        E = E + zy_sign(op.diag_zy[2 * d], op.diag_zy[2 * d + 1], s0, s1) * op.diag_coef[d];
    cx<T> dmine = mk(T(0), T(0));
    int tmine = -1;
    for (int t = 0; t < op.n_terms; t++) {
        const int fb = ldc(op.flip_begin, t), fe = ldc(op.flip_begin, t + 1);
        cx<T> v[JPL];
        {
            const int pos = ldc(op.flips, fb);
            const T f = spin_factor<T>(s0, s1, pos);
#pragma unroll
            for (int u = 0; u < JPL; u++) v[u] = th[u] + f * Wl[pos * ld + 16 * u];
        }
        for (int q = fb + 1; q < fe; q++) {
            const int pos = ldc(op.flips, q);
            const T f = spin_factor<T>(s0, s1, pos);
#pragma unroll
            for (int u = 0; u < JPL; u++) v[u] = v[u] + f * Wl[pos * ld + 16 * u];
        }
        cx<T> part = logcosh(v[0]);
#pragma unroll
        for (int u = 1; u < JPL; u++) part = part + logcosh(v[u]);
        cx<T> d = row_sum16(part - lsum);
        if (md.a) {
            for (int q = fb; q < fe; q++) {
                const int pos = op.flips[q];
                d = d + (T(-2) * spin_of<T>(s0, s1, pos)) * md.a[pos];
            }
        }
        if ((t & 15) == lane) { dmine = d; tmine = t; }
        if ((t & 15) == 15 || t == op.n_terms - 1) {
            if (tmine >= 0) E = E + term_coef(op, tmine, s0, s1) * cexp_wide(dmine);
            tmine = -1;
        }
    }
    return row_sum16(E);
}

template <typename T>
RBM_DI cx<T> visible_sum(const RbmDev<T>& md, uint64_t s0, uint64_t s1) {
    cx<T> v = mk(T(0), T(0));
    if (md.a)
        for (int i = 0; i < md.N; i++) v = v + spin_of<T>(s0, s1, i) * md.a[i];
    return v;
}

// theta_j = b_j + sum_i W_ij s_i  (include/quantum_state/Psi.hpp:43-53)
template <typename T, int JPL>
RBM_DI void init_angles(const RbmDev<T>& md, const cx<T>* __restrict__ Wsrc, uint64_t s0, uint64_t s1,
                        int lane, cx<T> (&th)[JPL]) {
    const int N = md.N, ld = md.ld;
#pragma unroll
    for (int u = 0; u < JPL; u++) th[u] = md.b[lane + 16 * u];
    for (int i = 0; i < N; i++) {
        const T si = spin_of<T>(s0, s1, i);
#pragma unroll
        for (int u = 0; u < JPL; u++) th[u] = th[u] + si * Wsrc[i * ld + lane + 16 * u];
    }
}

template <typename T, int JPL>
RBM_DI void measure(const RbmDev<T>& md, const cx<T>* __restrict__ Wsrc, const OpDev& op, int has_op,
                    uint64_t s0, uint64_t s1, const cx<T> (&th)[JPL], cx<T> vis, int lane, int64_t sample,
                    const Records& rec, bool valid = true) {
    const int M = md.M;
    // log psi from the cancellation-free form (logcosh_shifted) over the
    // model's hidden units, stored at once so that nothing of it stays live
    // through the E_loc loop
    {
        cx<T> lsum = lane < M ? logcosh_shifted(th[0]) : mk(T(0), T(0));
#pragma unroll
        for (int u = 1; u < JPL; u++)
            if (lane + 16 * u < M) lsum = lsum + logcosh_shifted(th[u]);
        const cx<double> lp = row_sum16(to_d(lsum)) + mk((double)M * LC_L0, 0.0) + to_d(vis);
        if (valid && lane == 0) {
            if (rec.spins) { rec.spins[2 * sample] = s0; rec.spins[2 * sample + 1] = s1; }
            if (rec.log_psi) rec.log_psi[sample] = lp;
        }
    }
    // E_loc's differences use the direct form summed over all slots (padded
    // ones included: they cancel in every difference)
    cx<T> lall = logcosh(th[0]);
#pragma unroll
    for (int u = 1; u < JPL; u++) lall = lall + logcosh(th[u]);
    cx<double> E = mk(0.0, 0.0);
    if (has_op) E = local_energy_rbm<T, JPL>(md, Wsrc, op, s0, s1, th, lall, lane);
    if (!valid) return;
    if (lane == 0 && rec.e_loc) rec.e_loc[sample] = E;
    if (rec.tanh) {
        cx<T>* Tr = reinterpret_cast<cx<T>*>(rec.tanh) + sample * (int64_t)M;
#pragma unroll
        for (int u = 0; u < JPL; u++) {
            const int j = lane + 16 * u;
            if (j < M) Tr[j] = dlogcosh(th[u]);
        }
    }
}

// ---------------------------------------------------------------------------
// Metropolis chains: include/spin_ensembles/MonteCarloLoop.hpp:55-256.
// Per call: random start, therm*N proposals, then for each of the
// steps_per_chain samples sweeps*N proposals followed by recording the
// configuration as sample index step * chains + chain.  Proposal k of a chain draws
// Philox(seed; chain, k >> 1, call, 0): words (x, y) for even k, (z, w) for
// odd k; position = x % N, u = (y + 0.5) 2^-32; accept iff ratio > 1 or
// u <= ratio with ratio = |psi(s')/psi(s)|^2.
//
// The draws do not depend on the chain state, so they are produced ahead of
// use: every 32 proposals the 16 lanes of a row compute the next 16 Philox
// blocks in parallel (lane l takes pair k/2 + l) and leave (position, y) in
// the row's slice of LDS.  The W row of the next proposal is loaded while the
// current one is evaluated.  Rows past the last chain replay the last chain
// and write nothing, so every wavefront reaches every barrier.
// ---------------------------------------------------------------------------
constexpr size_t MC_DRAW_LDS = 16 * MC_DRAW_BLOCK * sizeof(uint2);

// exp(2 delta), the ratio |psi(s')/psi(s)|^2
RBM_DI float ratio_of(float delta) { return __builtin_amdgcn_exp2f(delta * 2.8853900817779268f); }
RBM_DI double ratio_of(double delta) { return exp(2.0 * delta); }

// A draw in LDS is (position << 24 | position * ld, U) where U holds the
// uniform u = (y + 0.5) 2^-32: as FP32 bits for FP32 chains, as y itself for
// FP64 chains (which form u in FP64 like the FP64 restatement).
RBM_DI uint32_t pack_u(uint32_t y, float) { return __float_as_uint((float)(((double)y + 0.5) * (1.0 / 4294967296.0))); }
RBM_DI uint32_t pack_u(uint32_t y, double) { return y; }
RBM_DI float unpack_u(uint32_t v, float) { return __uint_as_float(v); }
RBM_DI double unpack_u(uint32_t v, double) { return ((double)v + 0.5) * (1.0 / 4294967296.0); }

// End of a chain: its last configuration (persistent chains) and its angles
// (drift diagnostics), when asked for.
template <typename T, int JPL>
RBM_DI void mc_chain_outputs(const RbmDev<T>& md, const McParams& mp, int chain, int lane, uint64_t s0, uint64_t s1,
                             const cx<T> (&th)[JPL]) {
    if (mp.final_out && lane == 0) {
        mp.final_out[2 * chain] = s0;
        mp.final_out[2 * chain + 1] = s1;
    }
    if (mp.angles_out) {
        cx<T>* out = reinterpret_cast<cx<T>*>(mp.angles_out) + (int64_t)chain * md.M;
#pragma unroll
        for (int u = 0; u < JPL; u++)
            if (lane + 16 * u < md.M) out[lane + 16 * u] = th[u];
    }
}

// NARROW: N <= 32, the configuration is one 32-bit word.
template <typename T, int JPL, bool WLDS, bool NARROW>
__global__ __launch_bounds__(256) void k_mc_rbm(RbmDev<T> md, McParams mp, uint64_t* __restrict__ spins_out,
                                                uint32_t* __restrict__ counts) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = md.N, ld = md.ld;
    const size_t wbytes = WLDS ? (size_t)N * ld * sizeof(cx<T>) : 0;
    const cx<T>* Wsrc = md.W;
    if (WLDS) {
        cx<T>* Wl = reinterpret_cast<cx<T>*>(smem_raw);
        for (int k = threadIdx.x; k < N * ld; k += blockDim.x) Wl[k] = md.W[k];
        Wsrc = Wl;
    }
    const int lane = threadIdx.x & 15, row = threadIdx.x >> 4;
    uint2* draws = reinterpret_cast<uint2*>(smem_raw + wbytes) + row * MC_DRAW_BLOCK;
    const int chain_raw = blockIdx.x * 16 + row;
    const bool valid = chain_raw < mp.chains;
    const int chain = valid ? chain_raw : mp.chains - 1;
    __syncthreads();

    uint64_t s0, s1;
    if (mp.start) {
        s0 = mp.start[2 * chain];
        s1 = mp.start[2 * chain + 1];
    } else {
        const u32x4 r0 = philox((uint32_t)chain, 0u, mp.call, 1u, mp.seed_lo, mp.seed_hi);
        s0 = (uint64_t)r0.x | ((uint64_t)r0.y << 32);
        s1 = (uint64_t)r0.z | ((uint64_t)r0.w << 32);
    }
    mask_words(s0, s1, N);
    uint32_t w32 = (uint32_t)s0;

    cx<T> th[JPL];
    init_angles<T, JPL>(md, Wsrc, s0, s1, lane, th);
    // Re lc(theta) summed over this lane's units (padded units included: they
    // stay at theta = 0 and cancel in every difference); a proposal's
    // log-ratio is the row sum of (the same sum at theta') - lane_lc.
    T lane_lc = logcosh_re(th[0]);
#pragma unroll
    for (int u = 1; u < JPL; u++) lane_lc += logcosh_re(th[u]);
    const cx<T>* __restrict__ Wl = Wsrc + lane;

    const uint64_t n_therm = (uint64_t)mp.therm * N;
    const uint64_t n_sweep = (uint64_t)mp.sweeps * N;
    const uint64_t total = n_therm + (uint64_t)mp.steps_per_chain * n_sweep;
    uint64_t next_measure = n_therm + n_sweep;
    // FP32 angle re-synchronisation point (checked at draw-block ends)
    const bool resync = sizeof(T) == 4 && mp.resync > 0;
    uint64_t next_resync = resync ? (uint64_t)mp.resync : ~(uint64_t)0;
    int sample = 0;
    uint32_t nacc = 0;

    // A sample is recorded as its configuration; log psi, E_loc and O_k are
    // evaluated afterwards by k_eval_rbm at full occupancy.
    auto measure_now = [&]() {
        if (valid && lane == 0) {
            const int64_t idx = (int64_t)sample * mp.chains + chain;
            spins_out[2 * idx] = NARROW ? (uint64_t)w32 : s0;
            spins_out[2 * idx + 1] = NARROW ? 0u : s1;
        }
        sample++;
        next_measure += n_sweep;
    };
    if (n_sweep == 0) {
        // num_sweeps = 0: every sample is taken right after thermalisation
        next_measure = ~(uint64_t)0;
    }
    if (n_sweep == 0 && n_therm == 0)
        while (sample < mp.steps_per_chain) measure_now();

    auto load_row = [&](cx<T> (&w)[JPL], uint32_t dx) {
#pragma unroll
        for (int u = 0; u < JPL; u++) w[u] = Wl[(dx & 0x7FFFu) + 16 * u];
    };
    // One proposal with draw d = (x, U) and the W row of its site in wr.
    auto propose = [&](uint2 d, const cx<T> (&wr)[JPL]) {
        const int p = (int)((d.x >> 16) & 127u);
        const T f = NARROW ? flip_factor_t<T>(w32 << (d.x >> 24)) : T(-2) * spin_of<T>(s0, s1, p);
        cx<T> thn[JPL];
        thn[0] = th[0] + f * wr[0];
        T lnew = logcosh_re(thn[0]);
#pragma unroll
        for (int u = 1; u < JPL; u++) {
            thn[u] = th[u] + f * wr[u];
            lnew += logcosh_re(thn[u]);
        }
        T delta = row_sum16(lnew - lane_lc);
        if (md.a) delta += f * md.a[p].re;
        // u < 1 always, so (ratio > 1 || u <= ratio) is u <= ratio
        const bool accept = unpack_u(d.y, T(0)) <= ratio_of(delta);
#pragma unroll
        for (int u = 0; u < JPL; u++) th[u] = accept ? thn[u] : th[u];
        lane_lc = accept ? lnew : lane_lc;
        if (NARROW) {
            w32 ^= (uint32_t)accept << p;
        } else {
            const uint64_t bit = (uint64_t)accept;
            s0 ^= p < 64 ? bit << (p & 63) : 0u;
            s1 ^= p >= 64 ? bit << (p & 63) : 0u;
        }
        nacc += accept ? 1u : 0u;
    };

    for (uint64_t kb = 0; kb < total; kb += MC_DRAW_BLOCK) {
        __syncthreads();
        {
            // x = position * ld | position << 16 | (31 - position) << 24
            const u32x4 r = philox((uint32_t)chain, (uint32_t)(((mp.k0 + kb) >> 1) + lane), mp.call, 0u, mp.seed_lo,
                                   mp.seed_hi);
            const uint32_t p0 = r.x % (uint32_t)N, p1 = r.z % (uint32_t)N;
            draws[2 * lane] = make_uint2(p0 * (uint32_t)ld | p0 << 16 | ((31u - p0) & 31u) << 24, pack_u(r.y, T(0)));
            draws[2 * lane + 1] = make_uint2(p1 * (uint32_t)ld | p1 << 16 | ((31u - p1) & 31u) << 24, pack_u(r.w, T(0)));
        }
        __syncthreads();
        const int nk = (int)(total - kb < (uint64_t)MC_DRAW_BLOCK ? total - kb : (uint64_t)MC_DRAW_BLOCK);
        int i = 0;
        while (i < nk) {
            // proposals up to the next measurement point or the end of the block;
            // the W row of the next proposal is loaded while the current one is
            // evaluated, alternating between two register sets, and the draw
            // after that is read one proposal ahead, so that no row load waits
            // on the LDS read that gives its address (at one wavefront per
            // SIMD nothing else would hide that latency)
            const int stop = next_measure > kb && next_measure - kb < (uint64_t)nk ? (int)(next_measure - kb) : nk;
            uint2 d = draws[i];
            uint2 dn = draws[i + 1 < nk ? i + 1 : i];
            cx<T> wa[JPL], wb[JPL];
            load_row(wa, d.x);
            while (i < stop) {
                // (the scheduling barriers keep these loads ahead of the
                // proposal instead of next to their first use)
                load_row(wb, dn.x);
                uint2 dnn = draws[i + 2 < nk ? i + 2 : nk - 1];
                __builtin_amdgcn_sched_barrier(0);
                propose(d, wa);
                d = dn;
                dn = dnn;
                if (++i >= stop) break;
                load_row(wa, dn.x);
                dnn = draws[i + 2 < nk ? i + 2 : nk - 1];
                __builtin_amdgcn_sched_barrier(0);
                propose(d, wb);
                d = dn;
                dn = dnn;
                ++i;
            }
            if (NARROW) s0 = w32;
            if (kb + (uint64_t)i == next_measure) measure_now();
            if (n_sweep == 0 && kb + (uint64_t)i == n_therm)
                while (sample < mp.steps_per_chain) measure_now();
        }
        if (resync && kb + (uint64_t)nk >= next_resync) {
            // recompute the angles from the configuration: the incremental
            // FP32 updates since the last re-sync leave their rounding behind
            init_angles<T, JPL>(md, Wsrc, s0, s1, lane, th);
            lane_lc = logcosh_re(th[0]);
#pragma unroll
            for (int u = 1; u < JPL; u++) lane_lc += logcosh_re(th[u]);
            next_resync += (uint64_t)mp.resync;
        }
    }
    if (lane == 0 && counts && valid) {
        counts[2 * chain] = nacc;
        counts[2 * chain + 1] = (uint32_t)(total - nacc);
    }
    if (valid) mc_chain_outputs<T, JPL>(md, mp, chain, lane, NARROW ? (uint64_t)w32 : s0, NARROW ? 0u : s1, th);
}

// ---------------------------------------------------------------------------
// Evaluation of given configurations (ExactSummation::kernel_foreach,
// include/spin_ensembles/ExactSummation.hpp:36-84, and the batched entry
// points).  Configuration idx is configs[idx] or, without a list,
// index_base + idx masked to N bits (Spins(index, N), Spins.h:230-237).
// Grid-strides over groups of 16 configurations per workgroup.
// ---------------------------------------------------------------------------
// Rows of W staged per step when W does not fit in LDS (FP32 models): the
// angles of a group's 16 configurations add W rows from LDS instead of each
// lane loading its units of every row from memory (100 dependent round trips
// per configuration at N = 100).
constexpr int EV_ROWS = 16;

template <typename T, int JPL, bool WLDS>
__global__ __launch_bounds__(256) void k_eval_rbm(RbmDev<T> md, OpDev op, int has_op, int64_t n,
                                                  const uint64_t* __restrict__ configs, uint64_t index_base,
                                                  Records rec) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = md.N, ld = md.ld;
    const cx<T>* Wsrc = md.W;
    if (WLDS) {
        cx<T>* Wl = reinterpret_cast<cx<T>*>(smem_raw);
        for (int k = threadIdx.x; k < N * ld; k += blockDim.x) Wl[k] = md.W[k];
        __syncthreads();
        Wsrc = Wl;
    }
    const int lane = threadIdx.x & 15;
    const int groups = blockDim.x >> 4;
    if constexpr (!WLDS && sizeof(T) == 4) {
        // theta = b + sum_i s_i W_i in the order of init_angles (the same FP32
        // values), the rows staged EV_ROWS at a time; every row of the
        // workgroup runs the same group loop, and rows past n evaluate the
        // last configuration and write nothing
        cx<T>* Wc = reinterpret_cast<cx<T>*>(smem_raw);
        const int row = threadIdx.x >> 4;
        for (int64_t g0 = (int64_t)blockIdx.x * groups; g0 < n; g0 += (int64_t)gridDim.x * groups) {
            const int64_t idx = g0 + row;
            const bool valid = idx < n;
            const int64_t ci = valid ? idx : n - 1;
            uint64_t s0, s1;
            if (configs) { s0 = configs[2 * ci]; s1 = configs[2 * ci + 1]; }
            else { s0 = index_base + (uint64_t)ci; s1 = 0; }
            mask_words(s0, s1, N);
            cx<T> th[JPL];
#pragma unroll
            for (int u = 0; u < JPL; u++) th[u] = md.b[lane + 16 * u];
            for (int i0 = 0; i0 < N; i0 += EV_ROWS) {
                const int nr = N - i0 < EV_ROWS ? N - i0 : EV_ROWS;
                __syncthreads();   // the previous rows have been used
                for (int k = threadIdx.x; k < nr * ld; k += blockDim.x) Wc[k] = md.W[(int64_t)i0 * ld + k];
                __syncthreads();
                for (int r = 0; r < nr; r++) {
                    const T si = spin_of<T>(s0, s1, i0 + r);
#pragma unroll
                    for (int u = 0; u < JPL; u++) th[u] = th[u] + si * Wc[r * ld + lane + 16 * u];
                }
            }
            const cx<T> vis = visible_sum(md, s0, s1);
            measure<T, JPL>(md, md.W, op, has_op, s0, s1, th, vis, lane, ci, rec, valid);
        }
        return;
    }
    for (int64_t idx = (int64_t)blockIdx.x * groups + (threadIdx.x >> 4); idx < n;
         idx += (int64_t)gridDim.x * groups) {
        uint64_t s0, s1;
        if (configs) { s0 = configs[2 * idx]; s1 = configs[2 * idx + 1]; }
        else { s0 = index_base + (uint64_t)idx; s1 = 0; }
        mask_words(s0, s1, N);
        cx<T> th[JPL];
        init_angles<T, JPL>(md, Wsrc, s0, s1, lane, th);
        const cx<T> vis = visible_sum(md, s0, s1);
        measure<T, JPL>(md, Wsrc, op, has_op, s0, s1, th, vis, lane, idx, rec);
    }
}

// ---------------------------------------------------------------------------
// Per-sample coefficients and scalar records (FP64).
//   weight w = mult[s] for runs of equal samples, else
//            exact ? num_configs * exp(2 (log prefactor + Re log psi)) : 1
//     (ExactSummation.hpp:75, Psi.hpp:133-136)
//   mode 0 (energy):   coef0 = w, coef1 = w E;  sc = {w, w E, w |E|^2}
//   mode 1 (distance): HilbertSpaceDistance.cu:83-105 with log psi' in lp2:
//     coef0 = w omega, coef1 = w ratio; sc = {w omega, w ratio, w norm}
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_coef(CoefArgs a) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < a.n; s += (int64_t)gridDim.x * blockDim.x) {
        const cx<double> lp = a.lp[s];
        const double w = a.mult ? a.mult[s] : a.exact ? a.num_configs * exp(2.0 * (a.log_prefactor + lp.re)) : 1.0;
        const cx<double> E = a.el ? a.el[s] : mk(0.0, 0.0);
        if (a.mode == 4) {
            // the weight itself: sum_s coef conj(O_k) = n <conj O_k>
            a.coef[s] = mk(w, 0.0);
            continue;
        }
        if (a.mode == 5) {
            a.coef[s] = w * (a.alpha * E + a.e_mean);
            continue;
        }
        if (a.mode == 6) {
            // ExpectationValue.cu:100-110: X = exp(2 Im E + Re(E^2 - E))
            double* sc = a.sc + s * SC_STRIDE;
            const cx<double> q = E * E - E;
            sc[0] = w * E.re; sc[1] = w * E.im; sc[2] = w * exp(2.0 * E.im + q.re); sc[3] = 0.0;
            continue;
        }
        if (a.mode == 7 || a.mode == 8) {
            // ExpectationValue.cu:392-412: r = |psi / psi'|^2 on samples of psi
            const double r = exp(2.0 * (lp.re - a.lp2[s].re));
            double* sc = a.sc + s * SC_STRIDE;
            if (a.mode == 8) {
                sc[0] = w * r; sc[1] = 0.0; sc[2] = 0.0; sc[3] = 0.0;
            } else {
                const cx<double> Ep = a.el2[s];
                sc[0] = w * Ep.re; sc[1] = w * Ep.im; sc[2] = w * r * E.re; sc[3] = w * r * E.im;
            }
            continue;
        }
        if (a.mode == 0 || a.mode == 2) {
            if (a.mode == 2) {
                // w (E - <E>): sum_s coef conj(O_k) = n (<E conj O_k> - <E> conj<O_k>)
                a.coef[s] = w * (E - a.e_mean);
                continue;
            }
            double* sc = a.sc + s * SC_STRIDE;
            sc[0] = w; sc[1] = w * E.re; sc[2] = w * E.im; sc[3] = w * (E.re * E.re + E.im * E.im);
        } else {
            const cx<double> lpp = a.lp2[s];
            const cx<double> dl = conj(lpp - lp);
            cx<double> omega;
            double nrm;
            if (a.is_unitary) {
                omega = cexp(dl) * E;
                nrm = E.re * E.re + E.im * E.im;
            } else {
                omega = cexp(E + dl);
                nrm = exp(2.0 * E.re);
            }
            const double ratio = exp(2.0 * (lpp.re - lp.re));
            if (a.mode == 3) {
                // w (r_mean omega - omega_mean r): the numerator of the distance
                // gradient in one coefficient (HilbertSpaceDistance.cu:257-262)
                a.coef[s] = w * (a.ratio_mean * omega - ratio * a.omega_mean);
                continue;
            }
            double* sc = a.sc + s * SC_STRIDE;
            sc[0] = w * omega.re; sc[1] = w * omega.im; sc[2] = w * ratio; sc[3] = w * nrm;
        }
    }
}

// E_k of the reference's gradient estimator (Operator.hpp:145-179): every
// string n maps s to s'_n with matrix element c_n(s), and
// E_k(s) = sum_n c_n(s) psi(s'_n)/psi(s) O_k(s'_n).  In canonical form the
// strings are the diagonal part and one entry per flip mask.  Sample i of the
// chunk [i0, i0 + cnt) expands to T + 1 virtual configurations: slot t < T is
// flip mask t with its matrix element, slot T is s itself with the diagonal
// matrix element.  vspins and vfac hold cnt * (T + 1) entries.
__global__ __launch_bounds__(256) void k_ek_expand(OpDev op, int64_t i0, int64_t cnt, const uint64_t* __restrict__ spins,
                                                   uint64_t* __restrict__ vspins, cx<double>* __restrict__ vfac) {
    const int T1 = op.n_terms + 1;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < cnt * T1; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + v / T1;
        const int t = (int)(v % T1);
        uint64_t s0 = spins[2 * i], s1 = spins[2 * i + 1];
        cx<double> f;
        if (t == op.n_terms) {
            f = diag_energy(op, s0, s1);
        } else {
            f = term_coef_lane(op, t, s0, s1);
            for (int q = op.flip_begin[t]; q < op.flip_begin[t + 1]; q++) flip_bit(s0, s1, op.flips[q]);
        }
        vspins[2 * v] = s0;
        vspins[2 * v + 1] = s1;
        vfac[v] = f;
    }
}

// Contraction coefficient of virtual configuration v of parent i:
// w_i (gamma E_i + delta) conj(c_v exp(log psi(s'_v) - log psi(s_i))), so
// that sum_v coef_v conj(O_k(s'_v)) = sum_i w_i (gamma E_i + delta) conj(E_k(s_i)).
// w holds the parents' weights (real parts); el may be null when gamma = 0.
__global__ __launch_bounds__(256) void k_ek_coef(int64_t i0, int64_t cnt, int T1, const cx<double>* __restrict__ w,
                                                 const cx<double>* __restrict__ lp, const cx<double>* __restrict__ el,
                                                 cx<double> gamma, cx<double> delta, const cx<double>* __restrict__ vlp,
                                                 const cx<double>* __restrict__ vfac, cx<double>* __restrict__ vcoef) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < cnt * T1; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + v / T1;
        const cx<double> e = vfac[v] * cexp(vlp[v] - lp[i]);
        const cx<double> f = el ? gamma * el[i] + delta : delta;
        vcoef[v] = w[i].re * (f * conj(e));
    }
}

// ---- free quantum axis of HilbertSpaceDistance (HilbertSpaceDistance.cu:47-80,
// 109-122).  axis holds N values each of da = alpha' - alpha, db = beta' - beta,
// cos(alpha' + alpha), sin(alpha' + alpha).  Samples are configs[2s..2s+1], or
// index_base + s when configs is null (the basis states of an exact sum).
RBM_DI void sample_words(const uint64_t* configs, uint64_t index_base, int64_t s, uint64_t& s0, uint64_t& s1) {
    if (configs) { s0 = configs[2 * s]; s1 = configs[2 * s + 1]; }
    else { s0 = index_base + (uint64_t)s; s1 = 0; }
}

// out[(i - i0) n + s] = sample s with spin i flipped, i in [i0, i0 + cnt)
__global__ __launch_bounds__(256) void k_flip_configs(int64_t n, const uint64_t* __restrict__ configs, uint64_t index_base,
                                                      int i0, int cnt, uint64_t* __restrict__ out) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n * cnt; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = v % n;
        uint64_t s0, s1;
        sample_words(configs, index_base, s, s0, s1);
        flip_bit(s0, s1, i0 + (int)(v / n));
        out[2 * v] = s0;
        out[2 * v + 1] = s1;
    }
}

// O of the axis angles of spin i: alpha_i -> s_i r_i, beta_i -> i (r_i cos_i - s_i sin_i)
RBM_DI cx<double> axis_o_beta(cx<double> r, double sp, double ca, double sa) {
    return mk(-r.im * ca, r.re * ca - sp * sa);
}

// r[i n + s] holds log psi'(s^i) on entry and r_i = psi'(s^i) / psi'(s) on exit;
// lp[s] (log psi'(s)) gets sum_i [da_i s_i r_i + db_i O_beta_i].
__global__ __launch_bounds__(256) void k_axis_log_psi(int64_t n, int N, const uint64_t* __restrict__ configs,
                                                      uint64_t index_base, const double* __restrict__ axis,
                                                      cx<double>* __restrict__ lp, cx<double>* __restrict__ r) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        uint64_t s0, s1;
        sample_words(configs, index_base, s, s0, s1);
        const cx<double> lp0 = lp[s];
        cx<double> corr = mk(0.0, 0.0);
        for (int i = 0; i < N; i++) {
            const cx<double> ri = cexp(r[(int64_t)i * n + s] - lp0);
            r[(int64_t)i * n + s] = ri;
            const double sp = spin_of<double>(s0, s1, i);
            corr = corr + (axis[i] * sp) * ri + axis[N + i] * axis_o_beta(ri, sp, axis[2 * N + i], axis[3 * N + i]);
        }
        lp[s] = lp0 + corr;
    }
}

// partials[y][k] = sum over samples of slice y of coef[s] conj(O_k(s)), k < 2N;
// workgroup (i, y) produces k = i and k = N + i.
__global__ __launch_bounds__(256) void k_axis_contract(int64_t n, int N, const uint64_t* __restrict__ configs,
                                                       uint64_t index_base, const double* __restrict__ axis,
                                                       const cx<double>* __restrict__ r, const cx<double>* __restrict__ coef,
                                                       cx<double>* __restrict__ partials) {
    __shared__ double red[4][256];
    const int i = blockIdx.x;
    const double ca = axis[2 * N + i], sa = axis[3 * N + i];
    cx<double> ga = mk(0.0, 0.0), gb = mk(0.0, 0.0);
    for (int64_t s = blockIdx.y * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.y * blockDim.x) {
        uint64_t s0, s1;
        sample_words(configs, index_base, s, s0, s1);
        const double sp = spin_of<double>(s0, s1, i);
        const cx<double> ri = r[(int64_t)i * n + s];
        const cx<double> c = coef[s];
        ga = ga + c * conj(sp * ri);
        gb = gb + c * conj(axis_o_beta(ri, sp, ca, sa));
    }
    red[0][threadIdx.x] = ga.re; red[1][threadIdx.x] = ga.im;
    red[2][threadIdx.x] = gb.re; red[3][threadIdx.x] = gb.im;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[(int64_t)blockIdx.y * 2 * N + i] = mk(red[0][0], red[1][0]);
        partials[(int64_t)blockIdx.y * 2 * N + N + i] = mk(red[2][0], red[3][0]);
    }
}

// Block partial sums of the SC_STRIDE scalar records, fixed order.
__global__ __launch_bounds__(256) void k_scalar_partials(int64_t n, const double* __restrict__ sc,
                                                         double* __restrict__ partials) {
    __shared__ double red[SC_STRIDE][256];
    double acc[SC_STRIDE];
#pragma unroll
    for (int q = 0; q < SC_STRIDE; q++) acc[q] = 0.0;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
        for (int q = 0; q < SC_STRIDE; q++) acc[q] += sc[s * SC_STRIDE + q];
#pragma unroll
    for (int q = 0; q < SC_STRIDE; q++) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < SC_STRIDE; q++) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < SC_STRIDE) partials[blockIdx.x * SC_STRIDE + threadIdx.x] = red[threadIdx.x][0];
}

// Log-sum-exp partials of x_s = 2 (log prefactor + Re log psi_s): per block
// (max, sum exp(x - max)).  Used for the norm (PsiNorm.cu:13-44) without
// overflow.
__global__ __launch_bounds__(256) void k_lse_partials(int64_t n, const cx<double>* __restrict__ lp,
                                                      double log_prefactor, double* __restrict__ partials) {
    __shared__ double rm[256], rs[256];
    double m = -INFINITY, sum = 0.0;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        const double x = 2.0 * (log_prefactor + lp[s].re);
        if (x > m) { sum = sum * exp(m - x) + 1.0; m = x; }
        else sum += exp(x - m);
    }
    rm[threadIdx.x] = m; rs[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double m1 = rm[threadIdx.x], m2 = rm[threadIdx.x + w];
            const double s1 = rs[threadIdx.x], s2 = rs[threadIdx.x + w];
            const double mm = m1 > m2 ? m1 : m2;
            rm[threadIdx.x] = mm;
            rs[threadIdx.x] = (m1 == -INFINITY ? 0.0 : s1 * exp(m1 - mm)) + (m2 == -INFINITY ? 0.0 : s2 * exp(m2 - mm));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = rm[0]; partials[2 * blockIdx.x + 1] = rs[0]; }
}

// ---------------------------------------------------------------------------
// O_k contractions.  With rows r = {1, s_0 .. s_{N-1}} and columns
// c = {T_0 .. T_{M-1}, 1} (T_j = my_tanh(theta_j)):
//   R[q][r][c] = sum_s coef_q[s] * row_r(s) * conj(col_c(s))
// covers sum w conj(O_b), sum w conj(O_W), sum w O_a and sum w (c = M, r = 0).
// Each block reduces a contiguous chunk of samples in FP64 and writes its
// partials; k_partials_sum adds blocks in fixed order (deterministic).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_ok_reduce(int64_t n, int N, int M, const cx<T>* __restrict__ Trec,
                                                   const uint64_t* __restrict__ spins,
                                                   const cx<double>* __restrict__ coef, int nc, int64_t chunk,
                                                   cx<double>* __restrict__ partials) {
    constexpr int RB = OK_ROWS_PER_THREAD;
    const int R = N + 1, C = M + 1;
    const int nrb = (R + RB - 1) / RB;
    const int nwork = nrb * C;
    const int64_t s_begin = (int64_t)blockIdx.x * chunk;
    const int64_t s_end = s_begin + chunk < n ? s_begin + chunk : n;
    const int64_t plen = (int64_t)nc * R * C;
    for (int w0 = 0; w0 < nwork; w0 += blockDim.x) {
        const int w = w0 + threadIdx.x;
        if (w >= nwork) break;
        const int col = w % C, rb = w / C;
        cx<double> acc[2][RB];
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int r = 0; r < RB; r++) acc[q][r] = mk(0.0, 0.0);
        for (int64_t s = s_begin; s < s_end; s++) {
            const cx<double> tv = col < M ? conj(to_d(Trec[s * M + col])) : mk(1.0, 0.0);
            const uint64_t s0 = spins[2 * s], s1 = spins[2 * s + 1];
            const cx<double> y0 = coef[s] * tv;
            const cx<double> y1 = nc > 1 ? coef[n + s] * tv : mk(0.0, 0.0);
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int row = rb * RB + r;
                const double sg = row == 0 ? 1.0 : spin_of<double>(s0, s1, row - 1);
                acc[0][r] = acc[0][r] + sg * y0;
                acc[1][r] = acc[1][r] + sg * y1;
            }
        }
        for (int q = 0; q < nc; q++)
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int row = rb * RB + r;
                if (row < R) partials[(int64_t)blockIdx.x * plen + ((int64_t)q * R + row) * C + col] = acc[q][r];
            }
    }
}

// ---------------------------------------------------------------------------
// O_k contraction.  Real GEMM over the samples:  P[r][n] = sum_s A_r(s) B_n(s),
//   A = row_r(s), rows {1, s_0..s_{N-1}} (exact +-1 values),
//   n = 2j + p: B = (p ? Im : Re)(coef(s) conj(col_j(s))), columns {T_0..T_{M-1}, 1}
// so that R[r][j] = sum_s coef row_r conj(col_j) = P[r][2j] + i P[r][2j+1].
// From 32 spins up it runs on the bfloat16 matrix cores with FP32 arithmetic: A is exact in
// bfloat16 and each FP32 value of B is split exactly into three bfloat16
// parts (hi = bf16(B), mid = bf16(B - hi), lo = B - hi - mid), so every
// product is exact and v_mfma_f32_16x16x32_bf16 sums it in FP32.
// A K-step takes 32 samples: lane l feeds samples 8 (l/16) .. 8 (l/16) + 7 for
// tile row / column l % 16 (A: lane l holds A[l%16][8(l/16) + q], B:
// B[8(l/16) + q][l%16], D: D[4(l/16)+v][l%16]).  One wavefront owns all RT row
// tiles and CT column tiles of one sample chunk, accumulates the chunk in FP32
// and writes its partials; chunks are added in FP64.  The coefficients are
// centred (zero mean), so the FP32 chunk sums carry no large cancellation.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));
// bfloat16 MFMA operands and the exact three-part split of FP32 values
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

RBM_DI uint32_t bf16_pair(float lo, float hi) {
    const __bf16 a = (__bf16)lo, b = (__bf16)hi;
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}
// three-part split of two FP32 values into packed bfloat16 pairs
RBM_DI void split3(float x0, float x1, uint32_t& hi, uint32_t& mid, uint32_t& lo) {
    hi = bf16_pair(x0, x1);
    const float r0 = x0 - __uint_as_float(hi << 16), r1 = x1 - __uint_as_float(hi & 0xFFFF0000u);
    mid = bf16_pair(r0, r1);
    lo = bf16_pair(r0 - __uint_as_float(mid << 16), r1 - __uint_as_float(mid & 0xFFFF0000u));
}


// Spin bit-planes of 32-sample blocks: planes[blk * N + i] has bit k set iff
// spin i of sample 32 blk + k is up (samples past n read as down).  One
// wavefront covers two blocks: a ballot per spin index collects the 64
// samples' bits, and lane k keeps the ballot of index base + k, so the stores
// are coalesced.
__global__ __launch_bounds__(256) void k_spin_planes(int64_t n, int N, int64_t nblk,
                                                     const uint64_t* __restrict__ spins,
                                                     uint32_t* __restrict__ planes) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t s = wave * 64 + lane;
    uint64_t w0 = 0, w1 = 0;
    if (s < n) { w0 = spins[2 * s]; w1 = spins[2 * s + 1]; }
    const int64_t blk = 2 * wave;
    for (int base = 0; base < N; base += 64) {
        uint64_t mine = 0;
        const int cnt = N - base < 64 ? N - base : 64;
        for (int k = 0; k < cnt; k++) {
            const int i = base + k;
            const uint64_t m = __ballot((int)((i < 64 ? w0 >> i : w1 >> (i - 64)) & 1u));
            mine = lane == k ? m : mine;
        }
        if (lane < cnt) {
            if (blk < nblk) planes[blk * N + base + lane] = (uint32_t)mine;
            if (blk + 1 < nblk) planes[(blk + 1) * N + base + lane] = (uint32_t)(mine >> 32);
        }
    }
}

// The O_k contraction on bfloat16 MFMA (v_mfma_f32_16x16x32_bf16, 32-sample
// K-steps) for models of at least 32 spins.  Operands of a K-step:
//   A (rows {1, s_0..s_{N-1}}, padding rows zero): exact +-1 bfloat16.  Lane
//     (ii, kk) of row tile a supplies row r = 16 a + ii for samples 8 kk..8 kk+7,
//     which is byte kk of the block's bit-plane of spin r - 1; a 256-entry
//     table in LDS turns that byte into the four packed bfloat16 pairs;
//   B (columns {Re, Im} x {coef conj T_j, coef}): every FP32 value split
//     exactly into three bfloat16 parts, so the three products are exact and
//     the MFMA sums them in FP32.
// RT row tiles (all rows) and CT column tiles per wavefront.  Workgroups are
// numbered so that all column groups of one sample chunk land on the same XCD
// (workgroup b runs on XCD b mod 8) and start together: they read the same
// my_tanh records at about the same time, so the repeats come from that XCD's
// L2 instead of HBM.  Chunks are whole 32-sample blocks.
template <int RT, int CT>
__global__ __launch_bounds__(64) void k_ok_mfma_bf3(int64_t n, int N, int M, const float* __restrict__ Tf,
                                                const uint32_t* __restrict__ planes,
                                                const cx<double>* __restrict__ coef, int64_t chunk, int nchunks,
                                                int ry, int cz, int rows_p, int cols_p, float* __restrict__ partials) {
    __shared__ uint4 lut[256];
    const int lane = threadIdx.x;
    // byte -> four bfloat16 pairs: sample 2q (low half) and 2q + 1 (high half)
    // of the byte are +1 (0x3F80) where their bit is set and -1 (0xBF80) else
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const uint32_t byte = (uint32_t)(lane + 64 * e);
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
            w[q] = 0xBF80BF80u ^ (((byte >> (2 * q)) & 1u) << 15) ^ (((byte >> (2 * q + 1)) & 1u) << 31);
        lut[byte] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    __syncthreads();
    const int kk = lane >> 4, ii = lane & 15;
    const int G = ry * cz;
    const int slot = blockIdx.x >> 3;
    const int chunk_id = (slot / G) * 8 + (blockIdx.x & 7);
    const int tile = slot % G;
    if (chunk_id >= nchunks) return;
    const int ct0 = (tile / ry) * CT;
    floatx4 acc[RT][CT];
#pragma unroll
    for (int a = 0; a < RT; a++)
#pragma unroll
        for (int b = 0; b < CT; b++) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    // column q = 2j + part: part = ii & 1 for every tile of this lane
    const int part = ii & 1;
    int coff[CT];
    float cfix[CT], cload[CT];
#pragma unroll
    for (int b = 0; b < CT; b++) {
        const int j = (16 * (ct0 + b) + ii) >> 1;
        coff[b] = j < M ? j : 0;                         // clamped: always a valid load
        cload[b] = j < M ? 1.f : 0.f;                    // take the loaded value
        cfix[b] = j == M ? 1.f : 0.f;                    // the constant column
    }
    // row kind of this lane in each row tile: 0 constant, 1 spin, 2 padding
    int rkind[RT], rspin[RT];
#pragma unroll
    for (int a = 0; a < RT; a++) {
        const int r = 16 * a + ii;
        rkind[a] = r == 0 ? 0 : (r <= N ? 1 : 2);
        rspin[a] = r >= 1 && r <= N ? r - 1 : 0;
    }
    const int64_t s_begin = (int64_t)chunk_id * chunk;
    const int64_t s_end = s_begin + chunk < n ? s_begin + chunk : n;
    const float2* T2 = reinterpret_cast<const float2*>(Tf);
    for (int64_t s0 = s_begin; s0 < s_end; s0 += 32) {
        const uint32_t* pl = planes + (s0 >> 5) * N;
        uint32_t bh[CT][4], bm[CT][4], bl[CT][4];
#pragma unroll
        for (int q2 = 0; q2 < 4; q2++) {
            float bv[2][CT];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int64_t s = s0 + 8 * kk + 2 * q2 + h;
                const bool ok = s < s_end;
                const int64_t sc = ok ? s : s_begin;
                const cx<double> c = coef[sc];
                const float cre = ok ? (float)c.re : 0.f, cim = ok ? (float)c.im : 0.f;
                // Re(c conj t) = cre tr + cim ti,  Im(c conj t) = cim tr - cre ti
                const float p = part ? cim : cre, qq = part ? -cre : cim;
                const float2* Tr = T2 + sc * M;
#pragma unroll
                for (int b = 0; b < CT; b++) {
                    const float2 t = Tr[coff[b]];
                    bv[h][b] = __builtin_fmaf(p, __builtin_fmaf(t.x, cload[b], cfix[b]), qq * (t.y * cload[b]));
                }
            }
#pragma unroll
            for (int b = 0; b < CT; b++) split3(bv[0][b], bv[1][b], bh[b][q2], bm[b][q2], bl[b][q2]);
        }
#pragma unroll
        for (int a = 0; a < RT; a++) {
            const uint32_t byte = rkind[a] == 0 ? 0xFFu : ((pl[rspin[a]] >> (8 * kk)) & 0xFFu);
            uint4 av = lut[byte];
            if (rkind[a] == 2) av = make_uint4(0u, 0u, 0u, 0u);
            const bf16x8_t A = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
            for (int b = 0; b < CT; b++) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    A, __builtin_bit_cast(bf16x8_t, make_uint4(bl[b][0], bl[b][1], bl[b][2], bl[b][3])), acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    A, __builtin_bit_cast(bf16x8_t, make_uint4(bm[b][0], bm[b][1], bm[b][2], bm[b][3])), acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    A, __builtin_bit_cast(bf16x8_t, make_uint4(bh[b][0], bh[b][1], bh[b][2], bh[b][3])), acc[a][b], 0, 0, 0);
            }
        }
    }
    float* out = partials + (int64_t)chunk_id * rows_p * cols_p;
#pragma unroll
    for (int a = 0; a < RT; a++)
#pragma unroll
        for (int b = 0; b < CT; b++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                out[(int64_t)(16 * a + 4 * kk + v) * cols_p + 16 * (ct0 + b) + ii] = acc[a][b][v];
}

// The same contraction on FP32 MFMA (v_mfma_f32_16x16x4f32, a K-step of 4
// samples; lane l holds A[l%16][l/16], B[l/16][l%16]) for models of at most 31
// spins: there the row tiles are few, and the bfloat16 kernel's per-sample
// column work (the exact split of 2(M+1) values) is not repaid by its faster
// matrix products.  RT x CT output tiles per wavefront.
template <int RT, int CT>
__global__ __launch_bounds__(64) void k_ok_mfma_f32(int64_t n, int N, int M, const float* __restrict__ Tf,
                                                const uint64_t* __restrict__ spins,
                                                const cx<double>* __restrict__ coef, int64_t chunk, int nchunks,
                                                int ry, int cz, int rows_p, int cols_p, float* __restrict__ partials) {
    const int lane = threadIdx.x;
    const int kk = lane >> 4, ii = lane & 15;
    const int G = ry * cz;
    const int slot = blockIdx.x >> 3;
    const int chunk_id = (slot / G) * 8 + (blockIdx.x & 7);
    const int tile = slot % G;
    if (chunk_id >= nchunks) return;
    const int rt0 = (tile % ry) * RT, ct0 = (tile / ry) * CT;
    floatx4 acc[RT][CT];
#pragma unroll
    for (int a = 0; a < RT; a++)
#pragma unroll
        for (int b = 0; b < CT; b++) acc[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    int rrow[RT], coff[CT];
    float cfix[CT], cload[CT];
#pragma unroll
    for (int a = 0; a < RT; a++) rrow[a] = 16 * (rt0 + a) + ii;
    // column q = 2j + part: part = ii & 1 for every tile of this lane
    const int part = ii & 1;
#pragma unroll
    for (int b = 0; b < CT; b++) {
        const int j = (16 * (ct0 + b) + ii) >> 1;
        coff[b] = j < M ? j : 0;                         // clamped: always a valid load
        cload[b] = j < M ? 1.f : 0.f;                    // take the loaded value
        cfix[b] = j == M ? 1.f : 0.f;                    // the constant column
    }
    const int64_t s_begin = (int64_t)chunk_id * chunk;
    const int64_t s_end = s_begin + chunk < n ? s_begin + chunk : n;
    const float2* T2 = reinterpret_cast<const float2*>(Tf);
#pragma unroll 2
    for (int64_t s0 = s_begin; s0 < s_end; s0 += 4) {
        const int64_t s = s0 + kk;
        const bool ok = s < s_end;
        const int64_t sc = ok ? s : s_begin;
        const cx<double> c = coef[sc];
        const uint64_t w0 = spins[2 * sc], w1 = spins[2 * sc + 1];
        const float cre = ok ? (float)c.re : 0.f, cim = ok ? (float)c.im : 0.f;
        // Re(c conj t) = cre tr + cim ti,  Im(c conj t) = cim tr - cre ti
        const float p = part ? cim : cre, q = part ? -cre : cim;
        const float2* Tr = T2 + sc * M;
        float bv[CT];
#pragma unroll
        for (int b = 0; b < CT; b++) {
            const float2 t = Tr[coff[b]];
            bv[b] = __builtin_fmaf(p, __builtin_fmaf(t.x, cload[b], cfix[b]), q * (t.y * cload[b]));
        }
        float av[RT];
#pragma unroll
        for (int a = 0; a < RT; a++) {
            const int r = rrow[a];
            av[a] = r == 0 ? 1.f : (r <= N ? spin_of<float>(w0, w1, (r - 1) & 127) : 0.f);
        }
#pragma unroll
        for (int a = 0; a < RT; a++)
#pragma unroll
            for (int b = 0; b < CT; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    float* out = partials + (int64_t)chunk_id * rows_p * cols_p;
#pragma unroll
    for (int a = 0; a < RT; a++)
#pragma unroll
        for (int b = 0; b < CT; b++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                out[(int64_t)(16 * (rt0 + a) + 4 * kk + v) * cols_p + 16 * (ct0 + b) + ii] = acc[a][b][v];
}

// Sum of FP32 chunk partials in FP64: block (x, y) adds chunks y, y + S, ... of
// 64 consecutive elements with 4 threads per element; out[y][len].
__global__ __launch_bounds__(256) void k_fsum(int nchunks, int S, int64_t len, const float* __restrict__ partials,
                                              double* __restrict__ out) {
    __shared__ double red[4][64];
    const int e_local = threadIdx.x & 63, sub = threadIdx.x >> 6;
    const int64_t e = (int64_t)blockIdx.x * 64 + e_local;
    double acc = 0.0;
    if (e < len)
        for (int b = blockIdx.y + S * sub; b < nchunks; b += 4 * S) acc += (double)partials[(int64_t)b * len + e];
    red[sub][e_local] = acc;
    __syncthreads();
    if (sub == 0 && e < len)
        out[(int64_t)blockIdx.y * len + e] = red[0][e_local] + red[1][e_local] + red[2][e_local] + red[3][e_local];
}

__global__ __launch_bounds__(256) void k_mfma_combine(int N, int M, int cols_p, int S, int64_t len,
                                                      const double* __restrict__ P, cx<double>* __restrict__ R) {
    const int C = M + 1;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < (N + 1) * C; e += gridDim.x * blockDim.x) {
        const int r = e / C, j = e % C;
        double re = 0, im = 0;
        for (int y = 0; y < S; y++) {
            const double* Py = P + (int64_t)y * len + (int64_t)r * cols_p + 2 * j;
            re += Py[0];
            im += Py[1];
        }
        R[e] = mk(re, im);
    }
}

// Tiling of the O_k contraction: rt = ceil((N + 1) / 16) row tiles and
// ct = ceil(2 (M + 1) / 16) column tiles.  FP32 kernel (rt <= 2): one or two
// row tiles and up to 13 column tiles per wavefront.  bfloat16 kernel: every
// row tile (N <= 128: at most 9) and two column tiles per wavefront.
static bool ok_uses_bf3(int N) { return (N + 1 + 15) / 16 > 2; }

static void mfma_tiling(int N, int M, int& RT, int& CT, int& rows_p, int& cols_p) {
    const int rt = (N + 1 + 15) / 16, ct = (2 * (M + 1) + 15) / 16;
    if (ok_uses_bf3(N)) {
        // column tiles per wavefront: the per-group work (row operands, the
        // sample loop) against the padding tiles, per 32 samples and in VALU
        // issue slots, with one MFMA counted as two
        RT = rt;
        CT = 2;
        int best = -1;
        for (int c = 2; c <= 4; c++) {
            const int groups = (ct + c - 1) / c;
            const int cost = groups * (4 * rt + 40) + groups * c * (76 + 6 * rt);
            if (best < 0 || cost <= best) { best = cost; CT = c; }
        }
    } else {
        RT = rt;
        static const int cts[] = {4, 8, 11, 13};
        const int groups = (ct + 12) / 13;
        const int need = (ct + groups - 1) / groups;
        CT = 13;
        for (int c : cts) if (c >= need) { CT = c; break; }
    }
    rows_p = 16 * RT;
    cols_p = 16 * CT * ((ct + CT - 1) / CT);
}

void ok_mfma_shape(int N, int M, int* rows_p, int* cols_p) {
    int RT, CT;
    mfma_tiling(N, M, RT, CT, *rows_p, *cols_p);
}

bool ok_mfma_needs_planes(int N) { return ok_uses_bf3(N); }

hipError_t launch_ok_mfma(int64_t n, int N, int M, const void* Trec, const uint64_t* spins, const cx<double>* coef,
                          int64_t chunk, int nchunks, int rows_p, int cols_p, float* partials, uint32_t* planes,
                          hipStream_t stream) {
    int RT, CT, rp, cp;
    mfma_tiling(N, M, RT, CT, rp, cp);
    if (rp != rows_p || cp != cols_p) return hipErrorInvalidValue;
    const int ry = rows_p / (16 * RT), cz = cols_p / (16 * CT);
    const dim3 grid((unsigned)(8 * ((nchunks + 7) / 8) * ry * cz));
    const float* Tf = reinterpret_cast<const float*>(Trec);
#define RBM_MF_CASE(K_, R_, C_, SP_)                                                                              \
    if (RT == R_ && CT == C_) {                                                                                    \
        hipLaunchKernelGGL((K_<R_, C_>), grid, dim3(64), 0, stream, n, N, M, Tf, SP_, coef, chunk, nchunks, ry, cz,   \
                           rows_p, cols_p, partials);                                                              \
        return hipGetLastError();                                                                                  \
    }
    if (ok_uses_bf3(N)) {
        // chunks are whole 32-sample blocks, and the bit-planes cover them
        if (!planes || chunk % 32) return hipErrorInvalidValue;
        const int64_t nblk = (n + 31) / 32;
        const int64_t waves = (nblk + 1) / 2;
        hipLaunchKernelGGL(k_spin_planes, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, stream, n, N, nblk, spins,
                           planes);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
#define RBM_MF_ROW(R_)                                                                                             \
    RBM_MF_CASE(k_ok_mfma_bf3, R_, 2, planes) RBM_MF_CASE(k_ok_mfma_bf3, R_, 3, planes)                            \
    RBM_MF_CASE(k_ok_mfma_bf3, R_, 4, planes)
        RBM_MF_ROW(3) RBM_MF_ROW(4) RBM_MF_ROW(5) RBM_MF_ROW(6) RBM_MF_ROW(7) RBM_MF_ROW(8) RBM_MF_ROW(9)
#undef RBM_MF_ROW
    } else {
#define RBM_MF_ROW(R_)                                                                                             \
    RBM_MF_CASE(k_ok_mfma_f32, R_, 4, spins) RBM_MF_CASE(k_ok_mfma_f32, R_, 8, spins)                              \
    RBM_MF_CASE(k_ok_mfma_f32, R_, 11, spins) RBM_MF_CASE(k_ok_mfma_f32, R_, 13, spins)
        RBM_MF_ROW(1) RBM_MF_ROW(2)
#undef RBM_MF_ROW
    }
#undef RBM_MF_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_mfma_finish(int nchunks, int N, int M, int rows_p, int cols_p, const float* partials, double* Psum,
                              cx<double>* R, hipStream_t stream) {
    const int64_t len = (int64_t)rows_p * cols_p;
    const int S = MFMA_SUM_SPLIT;
    hipLaunchKernelGGL(k_fsum, dim3((unsigned)((len + 63) / 64), (unsigned)S), dim3(256), 0, stream, nchunks, S, len,
                       partials, Psum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int outs = (N + 1) * (M + 1);
    hipLaunchKernelGGL(k_mfma_combine, dim3((unsigned)((outs + 255) / 256)), dim3(256), 0, stream, N, M, cols_p, S,
                       len, Psum, R);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_partials_sum(int nblocks, int64_t len, const cx<double>* __restrict__ partials,
                                                      cx<double>* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
        cx<double> s = mk(0.0, 0.0);
        for (int b = 0; b < nblocks; b++) s = s + partials[(int64_t)b * len + i];
        out[i] = s;
    }
}

// get_S_matrix (source/network_functions/S_matrix.cu:10-58, CPU semantics):
// S[k][k'] = inv * sum_s O_k(s) conj(O_k'(s)), unweighted and uncentred;
// rows/columns below 2N (alpha, beta) stay zero.
template <typename T>
__global__ __launch_bounds__(256) void k_s_matrix_ref(int64_t n, int N, int M, int has_a,
                                                      const cx<T>* __restrict__ Trec,
                                                      const uint64_t* __restrict__ spins, double inv,
                                                      cx<double>* __restrict__ S) {
    const int64_t P = 2 * (int64_t)N + M + (int64_t)N * M + (has_a ? N : 0);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < P * P; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = e / P - 2 * N, kp = e % P - 2 * N;
        cx<double> acc = mk(0.0, 0.0);
        if (k >= 0 && kp >= 0) {
            for (int64_t s = 0; s < n; s++) {
                const uint64_t s0 = spins[2 * s], s1 = spins[2 * s + 1];
                const cx<T>* Tr = Trec + s * M;
                auto ok = [&](int64_t kk) -> cx<double> {
                    if (kk < M) return to_d(Tr[kk]);
                    if (kk < M + (int64_t)N * M) {
                        const int64_t i = (kk - M) / M, j = (kk - M) % M;
                        return spin_of<double>(s0, s1, (int)i) * to_d(Tr[j]);
                    }
                    return mk(spin_of<double>(s0, s1, (int)(kk - M - (int64_t)N * M)), 0.0);
                };
                acc = acc + ok(k) * conj(ok(kp));
            }
            acc = inv * acc;
        }
        S[e] = acc;
    }
}

__global__ void k_activation(int64_t n, const cx<double>* __restrict__ z, cx<double>* __restrict__ out, int deriv) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = deriv ? dlogcosh(z[i]) : logcosh(z[i]);
}

// ---------------------------------------------------------------------------
// Host-side launchers.  JPL (hidden units per lane) is a compile-time
// parameter so theta stays in registers; dispatch picks the smallest JPL with
// 16 * JPL >= M.
// ---------------------------------------------------------------------------
#define RBM_JPL_LIST(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12) X(13) X(16)

static int pick_jpl(int M) {
    static const int list[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 13, 16};
    for (int j : list) if (16 * j >= M) return j;
    return -1;
}

// Metropolis chains with the exchange (total S_z) or fast-sweep move
// (MonteCarloLoop.hpp:153-256 with the move draws of propose_move).  One
// 16-lane row per chain as in k_mc_rbm; theta' adds the W rows of every spin
// that the move flips.  The chain records its configurations only.
template <typename T, int JPL>
__global__ __launch_bounds__(256) void k_mc_rbm_moves(RbmDev<T> md, McParams mp, uint64_t* __restrict__ spins_out,
                                                      uint32_t* __restrict__ counts) {
    const int N = md.N, ld = md.ld;
    const int lane = threadIdx.x & 15, row = threadIdx.x >> 4;
    const int chain_raw = blockIdx.x * 16 + row;
    const bool valid = chain_raw < mp.chains;
    const uint32_t chain = (uint32_t)(valid ? chain_raw : mp.chains - 1);
    uint64_t s0, s1;
    if (mp.start) {
        // a persistent chain continues from its last configuration (already in
        // its sector for the exchange move)
        s0 = mp.start[2 * chain];
        s1 = mp.start[2 * chain + 1];
        mask_words(s0, s1, N);
    } else {
        const u32x4 r0 = philox(chain, 0u, mp.call, 1u, mp.seed_lo, mp.seed_hi);
        s0 = (uint64_t)r0.x | ((uint64_t)r0.y << 32);
        s1 = (uint64_t)r0.z | ((uint64_t)r0.w << 32);
        mask_words(s0, s1, N);
        if (mp.move == MC_EXCHANGE) sector_start(N, mp.k_up, chain, mp.call, mp.seed_lo, mp.seed_hi, s0, s1);
    }
    cx<T> th[JPL];
    init_angles<T, JPL>(md, md.W, s0, s1, lane, th);
    T lane_lc = logcosh_re(th[0]);
#pragma unroll
    for (int u = 1; u < JPL; u++) lane_lc += logcosh_re(th[u]);
    const cx<T>* __restrict__ Wl = md.W + lane;
    const uint64_t n_therm = (uint64_t)mp.therm * N;
    const uint64_t n_sweep = (uint64_t)mp.sweeps * (mp.move == MC_FAST_SWEEP ? 1 : N);
    const uint64_t total = n_therm + (uint64_t)mp.steps_per_chain * n_sweep;
    uint64_t next_rec = n_therm + n_sweep;
    int step = 0;
    uint32_t nacc = 0;
    auto record = [&]() {
        if (valid && lane == 0) {
            const int64_t idx = (int64_t)step * mp.chains + chain;
            spins_out[2 * idx] = s0;
            spins_out[2 * idx + 1] = s1;
        }
        step++;
        next_rec += n_sweep;
    };
    while (step < mp.steps_per_chain && next_rec == 0) record();
    for (uint64_t k = 0; k < total; k++) {
        uint64_t n0, n1;
        uint32_t uw;
        propose_move(mp.move, mp.tries, N, chain, k, mp.call, mp.seed_lo, mp.seed_hi, s0, s1, n0, n1, uw);
        cx<T> thn[JPL];
#pragma unroll
        for (int u = 0; u < JPL; u++) thn[u] = th[u];
        T vis = T(0);
        uint64_t d0 = s0 ^ n0, d1 = s1 ^ n1;
        while (d0 | d1) {
            const int i = d0 ? __builtin_ctzll(d0) : 64 + __builtin_ctzll(d1);
            if (d0) d0 &= d0 - 1u; else d1 &= d1 - 1u;
            const T f = T(-2) * spin_of<T>(s0, s1, i);
#pragma unroll
            for (int u = 0; u < JPL; u++) thn[u] = thn[u] + f * Wl[i * ld + 16 * u];
            if (md.a) vis += f * md.a[i].re;
        }
        T lnew = logcosh_re(thn[0]);
#pragma unroll
        for (int u = 1; u < JPL; u++) lnew += logcosh_re(thn[u]);
        const T delta = row_sum16(lnew - lane_lc) + vis;
        const bool accept = unpack_u(pack_u(uw, T(0)), T(0)) <= ratio_of(delta);
        if (accept) {
#pragma unroll
            for (int u = 0; u < JPL; u++) th[u] = thn[u];
            lane_lc = lnew;
            s0 = n0; s1 = n1;
            nacc++;
        }
        while (step < mp.steps_per_chain && k + 1 == next_rec) record();
        if (sizeof(T) == 4 && mp.resync > 0 && (k + 1) % (uint64_t)mp.resync == 0) {
            init_angles<T, JPL>(md, md.W, s0, s1, lane, th);
            lane_lc = logcosh_re(th[0]);
#pragma unroll
            for (int u = 1; u < JPL; u++) lane_lc += logcosh_re(th[u]);
        }
    }
    if (lane == 0 && counts && valid) {
        counts[2 * chain] = nacc;
        counts[2 * chain + 1] = (uint32_t)(total - nacc);
    }
    if (valid) mc_chain_outputs<T, JPL>(md, mp, (int)chain, lane, s0, s1, th);
}

// log psi'(s) - log psi(s) for two RBMs of the same shape from the angle
// difference delta = (b' - b) + (W' - W)^T s, unit by unit with
// logcosh_diff, plus sum_i (a' - a)_i s_i in FP64; lp2 = lp + that difference.
// Subtracting two complex64 sums of M log-cosh values would lose the small
// difference of nearby wavefunctions (HilbertSpaceDistance.cu:40-44, 83-105).
template <typename T, int JPL>
__global__ __launch_bounds__(256) void k_rbm_log_ratio(RbmDev<T> md, const cx<T>* __restrict__ dW,
                                                       const cx<T>* __restrict__ db, const cx<double>* __restrict__ da,
                                                       int64_t n, const uint64_t* __restrict__ configs,
                                                       uint64_t index_base, const cx<double>* __restrict__ lp,
                                                       cx<double>* __restrict__ lp2) {
    const int N = md.N, ld = md.ld;
    const int lane = threadIdx.x & 15;
    const int groups = blockDim.x >> 4;
    for (int64_t idx = (int64_t)blockIdx.x * groups + (threadIdx.x >> 4); idx < n;
         idx += (int64_t)gridDim.x * groups) {
        uint64_t s0, s1;
        if (configs) { s0 = configs[2 * idx]; s1 = configs[2 * idx + 1]; }
        else { s0 = index_base + (uint64_t)idx; s1 = 0; }
        mask_words(s0, s1, N);
        cx<T> th[JPL], dl[JPL];
        init_angles<T, JPL>(md, md.W, s0, s1, lane, th);
#pragma unroll
        for (int u = 0; u < JPL; u++) dl[u] = db[lane + 16 * u];
        for (int i = 0; i < N; i++) {
            const T si = spin_of<T>(s0, s1, i);
#pragma unroll
            for (int u = 0; u < JPL; u++) dl[u] = dl[u] + si * dW[i * ld + lane + 16 * u];
        }
        cx<T> acc = logcosh_diff(th[0], dl[0]);
#pragma unroll
        for (int u = 1; u < JPL; u++) acc = acc + logcosh_diff(th[u], dl[u]);
        cx<double> d = row_sum16(to_d(acc));
        if (lane == 0) {
            if (da)
                for (int i = 0; i < N; i++) d = d + spin_of<double>(s0, s1, i) * da[i];
            lp2[idx] = lp[idx] + d;
        }
    }
}

template <typename T>
hipError_t launch_rbm_log_ratio(const RbmDev<T>& md, const void* dW, const void* db, const cx<double>* da, int64_t n,
                                const uint64_t* configs, uint64_t index_base, const cx<double>* lp, cx<double>* lp2,
                                int max_blocks, hipStream_t stream) {
    const int jpl = pick_jpl(md.M);
    int64_t blocks = (n + 15) / 16;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    const cx<T>* w = reinterpret_cast<const cx<T>*>(dW);
    const cx<T>* b = reinterpret_cast<const cx<T>*>(db);
#define RBM_RATIO_CASE(J)                                                                                           \
    case J:                                                                                                         \
        hipLaunchKernelGGL((k_rbm_log_ratio<T, J>), dim3((unsigned)blocks), dim3(256), 0, stream, md, w, b, da, n, \
                           configs, index_base, lp, lp2);                                                            \
        break;
    switch (jpl) {
        RBM_JPL_LIST(RBM_RATIO_CASE)
        default: return hipErrorInvalidValue;
    }
#undef RBM_RATIO_CASE
    return hipGetLastError();
}
template hipError_t launch_rbm_log_ratio<float>(const RbmDev<float>&, const void*, const void*, const cx<double>*, int64_t,
                                                const uint64_t*, uint64_t, const cx<double>*, cx<double>*, int, hipStream_t);
template hipError_t launch_rbm_log_ratio<double>(const RbmDev<double>&, const void*, const void*, const cx<double>*,
                                                 int64_t, const uint64_t*, uint64_t, const cx<double>*, cx<double>*, int,
                                                 hipStream_t);

int max_hidden_units() { return 16 * 16; }
int rbm_padded_width(int M) { return 16 * pick_jpl(M); }

template <typename T>
hipError_t launch_mc(const RbmDev<T>& md, const McParams& mp, uint64_t* spins, uint32_t* counts, bool wlds,
                     size_t lds_bytes, hipStream_t stream) {
    const int jpl = pick_jpl(md.M);
    const dim3 block(256), grid((mp.chains + 15) / 16);
    const size_t shm = (wlds ? lds_bytes : 0) + MC_DRAW_LDS;
    const bool narrow = md.N <= 32 && wlds;
#define RBM_MC_CASE(J)                                                                                          \
    case J:                                                                                                     \
        if (narrow) {                                                                                           \
            if (shm > 65536)                                                                                    \
                (void)hipFuncSetAttribute((const void*)&k_mc_rbm<T, J, true, true>,                             \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);                 \
            hipLaunchKernelGGL((k_mc_rbm<T, J, true, true>), grid, block, shm, stream, md, mp, spins, counts);  \
        } else if (wlds) {                                                                                      \
            if (shm > 65536)                                                                                    \
                (void)hipFuncSetAttribute((const void*)&k_mc_rbm<T, J, true, false>,                            \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);                 \
            hipLaunchKernelGGL((k_mc_rbm<T, J, true, false>), grid, block, shm, stream, md, mp, spins, counts); \
        } else                                                                                                  \
            hipLaunchKernelGGL((k_mc_rbm<T, J, false, false>), grid, block, shm, stream, md, mp, spins, counts); \
        break;
    switch (jpl) {
        RBM_JPL_LIST(RBM_MC_CASE)
        default: return hipErrorInvalidValue;
    }
#undef RBM_MC_CASE
    return hipGetLastError();
}

template <typename T>
hipError_t launch_eval(const RbmDev<T>& md, const OpDev& op, int has_op, int64_t n, const uint64_t* configs,
                       uint64_t index_base, const Records& rec, bool wlds, size_t lds_bytes, int max_blocks,
                       hipStream_t stream) {
    const int jpl = pick_jpl(md.M);
    int64_t blocks = (n + 15) / 16;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    const dim3 block(256), grid((unsigned)blocks);
    const size_t shm = wlds ? lds_bytes : 0;
    // FP32 models whose W is not in LDS stage EV_ROWS rows of it at a time
    const size_t rows_shm = sizeof(T) == 4 ? (size_t)EV_ROWS * md.ld * sizeof(cx<T>) : 0;
#define RBM_EVAL_CASE(J)                                                                                              \
    case J:                                                                                                           \
        if (wlds) {                                                                                                   \
            if (shm > 65536)                                                                                          \
                (void)hipFuncSetAttribute((const void*)&k_eval_rbm<T, J, true>,                                       \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);                       \
            hipLaunchKernelGGL((k_eval_rbm<T, J, true>), grid, block, shm, stream, md, op, has_op, n, configs, index_base, rec); \
        }                                                                                                             \
        else hipLaunchKernelGGL((k_eval_rbm<T, J, false>), grid, block, rows_shm, stream, md, op, has_op, n, configs, index_base, rec); \
        break;
    switch (jpl) {
        RBM_JPL_LIST(RBM_EVAL_CASE)
        default: return hipErrorInvalidValue;
    }
#undef RBM_EVAL_CASE
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ok_reduce(int64_t n, int N, int M, const void* Trec, const uint64_t* spins, const cx<double>* coef,
                            int nc, int64_t chunk, int nblocks, cx<double>* partials, hipStream_t stream) {
    hipLaunchKernelGGL((k_ok_reduce<T>), dim3(nblocks), dim3(256), 0, stream, n, N, M,
                       reinterpret_cast<const cx<T>*>(Trec), spins, coef, nc, chunk, partials);
    return hipGetLastError();
}

hipError_t launch_partials_sum(int nblocks, int64_t len, const cx<double>* partials, cx<double>* out, hipStream_t stream) {
    int64_t g = (len + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_partials_sum, dim3((unsigned)g), dim3(256), 0, stream, nblocks, len, partials, out);
    return hipGetLastError();
}

hipError_t launch_coef(const CoefArgs& a, hipStream_t stream) {
    int64_t g = (a.n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_coef, dim3((unsigned)g), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_ek_expand(const OpDev& op, int64_t i0, int64_t cnt, const uint64_t* spins, uint64_t* vspins,
                            cx<double>* vfac, hipStream_t stream) {
    int64_t g = (cnt * (op.n_terms + 1) + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_ek_expand, dim3((unsigned)g), dim3(256), 0, stream, op, i0, cnt, spins, vspins, vfac);
    return hipGetLastError();
}

hipError_t launch_ek_coef(int64_t i0, int64_t cnt, int T1, const cx<double>* w, const cx<double>* lp, const cx<double>* el,
                          cx<double> gamma, cx<double> delta, const cx<double>* vlp, const cx<double>* vfac,
                          cx<double>* vcoef, hipStream_t stream) {
    int64_t g = (cnt * T1 + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_ek_coef, dim3((unsigned)g), dim3(256), 0, stream, i0, cnt, T1, w, lp, el, gamma, delta, vlp, vfac,
                       vcoef);
    return hipGetLastError();
}

hipError_t launch_flip_configs(int64_t n, const uint64_t* configs, uint64_t index_base, int i0, int cnt, uint64_t* out,
                               hipStream_t stream) {
    int64_t g = (n * cnt + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_flip_configs, dim3((unsigned)g), dim3(256), 0, stream, n, configs, index_base, i0, cnt, out);
    return hipGetLastError();
}

hipError_t launch_axis_log_psi(int64_t n, int N, const uint64_t* configs, uint64_t index_base, const double* axis,
                               cx<double>* lp, cx<double>* r, hipStream_t stream) {
    int64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_axis_log_psi, dim3((unsigned)g), dim3(256), 0, stream, n, N, configs, index_base, axis, lp, r);
    return hipGetLastError();
}

hipError_t launch_axis_contract(int64_t n, int N, const uint64_t* configs, uint64_t index_base, const double* axis,
                                const cx<double>* r, const cx<double>* coef, int slices, cx<double>* partials,
                                hipStream_t stream) {
    hipLaunchKernelGGL(k_axis_contract, dim3((unsigned)N, (unsigned)slices), dim3(256), 0, stream, n, N, configs, index_base,
                       axis, r, coef, partials);
    return hipGetLastError();
}

hipError_t launch_scalar_partials(int64_t n, const double* sc, int nblocks, double* partials, hipStream_t stream) {
    hipLaunchKernelGGL(k_scalar_partials, dim3(nblocks), dim3(256), 0, stream, n, sc, partials);
    return hipGetLastError();
}

hipError_t launch_lse_partials(int64_t n, const cx<double>* lp, double log_prefactor, int nblocks, double* partials,
                               hipStream_t stream) {
    hipLaunchKernelGGL(k_lse_partials, dim3(nblocks), dim3(256), 0, stream, n, lp, log_prefactor, partials);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_s_matrix_ref(int64_t n, int N, int M, int has_a, const void* Trec, const uint64_t* spins, double inv,
                               cx<double>* S, hipStream_t stream) {
    const int64_t P = 2 * (int64_t)N + M + (int64_t)N * M + (has_a ? N : 0);
    int64_t g = (P * P + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL((k_s_matrix_ref<T>), dim3((unsigned)g), dim3(256), 0, stream, n, N, M, has_a,
                       reinterpret_cast<const cx<T>*>(Trec), spins, inv, S);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// log psi of n configurations (RBM without visible bias), theta = b + W^T s on
// the bfloat16 matrix cores with FP32 arithmetic: the +-1 spins are exact in
// bfloat16 and every FP32 entry of W is split exactly into three bfloat16
// parts, W = hi + mid + lo (hi = bf16(W), mid = bf16(W - hi), lo = W - hi -
// mid), so each product s_k W_kj is exact and the MFMA sums it in FP32, as an
// FP32 GEMM does.  my_logcosh is summed over the units in the epilogue
// (Psi::log_psi_s, include/quantum_state/Psi.hpp:55-63).
// v_mfma_f32_16x16x32_bf16: lane l supplies A[l % 16][8 (l / 16) + q] (spin k =
// 32 kb + 8 (l / 16) + q of configuration row l % 16) and B[8 (l / 16) + q][l % 16]
// (W[k][unit]); accumulator element v of lane l is D[4 (l / 16) + v][l % 16].
// A workgroup of 8 waves takes 256 configurations (two 16-row tiles per wave)
// and stages the split W of 16 units at a time in LDS as B fragments (24 KiB).
// ---------------------------------------------------------------------------
constexpr int TL_WAVES = 8, TL_TILES = 2, TL_CT = 1, TL_UNITS = 16 * TL_CT, TL_KBMAX = 4;

__global__ __launch_bounds__(512) void k_theta_logpsi(RbmDev<float> md, int64_t n, const uint64_t* __restrict__ spins,
                                                      cx<double>* __restrict__ lp) {
    // B fragments of one chunk: [kb][column tile c][re, im][hi, mid, lo][lane]
    __shared__ uint4 Wf[TL_KBMAX][TL_CT][2][3][64];
    const int N = md.N, M = md.M, ld = md.ld;
    const int KB = (N + 31) >> 5;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int64_t base = ((int64_t)blockIdx.x * TL_WAVES + wave) * (16 * TL_TILES);
    // A fragments: 8 spins per K block as bfloat16 +-1 (0 past N)
    uint32_t af[TL_TILES][TL_KBMAX][4];
#pragma unroll
    for (int r = 0; r < TL_TILES; r++) {
        const int64_t p = base + 16 * r + (lane & 15);
        uint64_t s0 = 0, s1 = 0;
        if (p < n) { s0 = spins[2 * p]; s1 = spins[2 * p + 1]; }
#pragma unroll
        for (int kb = 0; kb < TL_KBMAX; kb++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k0 = 32 * kb + 8 * g + 2 * q;
                const uint32_t e0 = k0 < N ? (bit_of(s0, s1, k0) ? 0x3F80u : 0xBF80u) : 0u;
                const uint32_t e1 = k0 + 1 < N ? (bit_of(s0, s1, k0 + 1) ? 0x3F80u : 0xBF80u) : 0u;
                af[r][kb][q] = e0 | (e1 << 16);
            }
    }
    float lre[TL_TILES][4], lim[TL_TILES][4];
#pragma unroll
    for (int r = 0; r < TL_TILES; r++)
#pragma unroll
        for (int v = 0; v < 4; v++) { lre[r][v] = 0.0f; lim[r][v] = 0.0f; }
    for (int u0 = 0; u0 < ld; u0 += TL_UNITS) {
        __syncthreads();
        // fragment slot (kb, c, L): unit u0 + 16 c + L % 16, k = 32 kb + 8 (L / 16) + q
        for (int f = threadIdx.x; f < KB * TL_CT * 64; f += blockDim.x) {
            const int kb = f / (TL_CT * 64), c = (f >> 6) % TL_CT, L = f & 63;
            const int unit = u0 + 16 * c + (L & 15);
            const int kbase = 32 * kb + 8 * (L >> 4);
            uint32_t h[2][4], m[2][4], o[2][4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                cx<float> w0 = mk(0.0f, 0.0f), w1 = mk(0.0f, 0.0f);
                const int k0 = kbase + 2 * q;
                if (unit < ld) {
                    if (k0 < N) w0 = md.W[k0 * ld + unit];
                    if (k0 + 1 < N) w1 = md.W[(k0 + 1) * ld + unit];
                }
                split3(w0.re, w1.re, h[0][q], m[0][q], o[0][q]);
                split3(w0.im, w1.im, h[1][q], m[1][q], o[1][q]);
            }
#pragma unroll
            for (int ri = 0; ri < 2; ri++) {
                Wf[kb][c][ri][0][L] = make_uint4(h[ri][0], h[ri][1], h[ri][2], h[ri][3]);
                Wf[kb][c][ri][1][L] = make_uint4(m[ri][0], m[ri][1], m[ri][2], m[ri][3]);
                Wf[kb][c][ri][2][L] = make_uint4(o[ri][0], o[ri][1], o[ri][2], o[ri][3]);
            }
        }
        __syncthreads();
        floatx4 acc[TL_TILES][TL_CT][2];
#pragma unroll
        for (int r = 0; r < TL_TILES; r++)
#pragma unroll
            for (int c = 0; c < TL_CT; c++) {
                acc[r][c][0] = floatx4{0.f, 0.f, 0.f, 0.f};
                acc[r][c][1] = floatx4{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
        for (int kb = 0; kb < TL_KBMAX; kb++) {
            if (kb >= KB) break;
#pragma unroll
            for (int c = 0; c < TL_CT; c++)
#pragma unroll
                for (int ri = 0; ri < 2; ri++)
#pragma unroll
                    for (int part = 2; part >= 0; part--) {   // smallest parts first
                        const bf16x8_t B = __builtin_bit_cast(bf16x8_t, Wf[kb][c][ri][part][lane]);
#pragma unroll
                        for (int r = 0; r < TL_TILES; r++) {
                            const bf16x8_t A = __builtin_bit_cast(
                                bf16x8_t, make_uint4(af[r][kb][0], af[r][kb][1], af[r][kb][2], af[r][kb][3]));
                            acc[r][c][ri] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[r][c][ri], 0, 0, 0);
                        }
                    }
        }
#pragma unroll
        for (int c = 0; c < TL_CT; c++) {
            const int j = u0 + 16 * c + (lane & 15);
            if (j < M) {
                const cx<float> bj = md.b[j];
#pragma unroll
                for (int r = 0; r < TL_TILES; r++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        // the cancellation-free form: log psi is the result here
                        const cx<float> lc = logcosh_shifted(mk(acc[r][c][0][v] + bj.re, acc[r][c][1][v] + bj.im));
                        lre[r][v] += lc.re;
                        lim[r][v] += lc.im;
                    }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < TL_TILES; r++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const float sr = row_sum16(lre[r][v]);
            const float si = row_sum16(lim[r][v]);
            const int64_t p = base + 16 * r + 4 * g + v;
            if ((lane & 15) == 0 && p < n) lp[p] = mk((double)sr + (double)md.M * LC_L0, (double)si);
        }
}

hipError_t launch_theta_logpsi(const RbmDev<float>& md, int64_t n, const uint64_t* spins, cx<double>* lp,
                               hipStream_t stream) {
    if (md.a || md.N > 32 * TL_KBMAX || n <= 0) return hipErrorInvalidValue;
    const int64_t per_block = (int64_t)TL_WAVES * 16 * TL_TILES;
    const int64_t blocks = (n + per_block - 1) / per_block;
    hipLaunchKernelGGL(k_theta_logpsi, dim3((unsigned)blocks), dim3(64 * TL_WAVES), 0, stream, md, n, spins, lp);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Centred, weighted S-matrix S_kl = (1/n) sum_s w (O_k - <O_k>)^* (O_l - <O_l>)
// over the RBM parameters with an O_k (b, W, visible a; order of `params`
// after alpha and beta).  The rows of a sample batch are materialised
// "realified", [Re O (P) | Im O (P) | 0 padding] scaled by sqrt(w), so that one
// real product G = O^T O gives S_kl = G[k][l] + G[P+k][P+l] + i (G[k][P+l] - G[P+k][l]).
// O_k(s): b_j -> T_j, W_ij -> s_i T_j, a_i -> s_i  (Psi.hpp:165-202).
// Each FP32 value is split exactly into three bfloat16 parts here, once, for
// the Gram kernel's matrix cores, and stored transposed per block of 32
// samples: part q of column c and sample 32 b + k at
// out[((b * 3 + q) * ld2 + c) * 32 + k].
// ---------------------------------------------------------------------------
// One workgroup per (32-sample block blockIdx.y, 128 columns from
// 128 blockIdx.x).  The per-sample factors (sqrt w, spin words) are staged in
// LDS once and each thread decodes its column once; the values are formed
// with consecutive threads on consecutive columns (the my_tanh records of a
// sample are read along its row) into LDS, then split and written with
// consecutive threads on the 16 sample pairs of consecutive columns, so both
// the reads and the 3 x 8 KB of writes are contiguous.
template <typename T>
__global__ __launch_bounds__(256) void k_o_split3(int64_t s0, int64_t nb, int N, int M, int P, int ld2,
                                                  const void* __restrict__ Trec, const uint64_t* __restrict__ spins,
                                                  const cx<double>* __restrict__ w, const cx<float>* __restrict__ obar,
                                                  uint32_t* __restrict__ out) {
    __shared__ float vals[32][129];
    __shared__ float rsw[32];
    __shared__ uint64_t rsp[32][2];
    const cx<T>* Tr = reinterpret_cast<const cx<T>*>(Trec);
    const int64_t blk = blockIdx.y;
    if (threadIdx.x < 32) {
        const int64_t sr = blk * 32 + threadIdx.x;
        const bool in = sr < nb;
        rsw[threadIdx.x] = in ? (float)sqrt(w[s0 + sr].re) : 0.0f;
        rsp[threadIdx.x][0] = in ? spins[2 * (s0 + sr)] : 0;
        rsp[threadIdx.x][1] = in ? spins[2 * (s0 + sr) + 1] : 0;
    }
    const int c0 = blockIdx.x * 128;
    const int cc = threadIdx.x & 127;
    const int c = c0 + cc;
    // column c: O_k with k = c mod P, real part for c < P, imaginary part after
    const bool live = c < 2 * P;
    const bool re_part = c < P;
    const int k = live ? (re_part ? c : c - P) : 0;
    int j = 0, si_idx = -1;   // T_j factor (absent for the visible biases), spin factor s_i
    bool has_t = true;
    if (k < M) {
        j = k;
    } else if (k < M + N * M) {
        si_idx = (k - M) / M;
        j = (k - M) % M;
    } else {
        has_t = false;
        si_idx = k - M - N * M;
    }
    const cx<float> ob = obar[k];
    __syncthreads();
    for (int r = threadIdx.x >> 7; r < 32; r += 2) {
        const int64_t sr = blk * 32 + r;
        float v = 0.0f;
        if (live && sr < nb) {
            cx<float> o = mk(1.0f, 0.0f);
            if (has_t) {
                const cx<T> tj = Tr[(s0 + sr) * M + j];
                o = mk((float)tj.re, (float)tj.im);
            }
            if (si_idx >= 0) {
                const float si = spin_of<float>(rsp[r][0], rsp[r][1], si_idx);
                o = mk(si * o.re, si * o.im);
            }
            const cx<float> d = o - ob;
            v = rsw[r] * (re_part ? d.re : d.im);
        }
        vals[r][cc] = v;
    }
    __syncthreads();
    const int kp = threadIdx.x & 15;
    for (int cw = threadIdx.x >> 4; cw < 128; cw += 16) {
        uint32_t hi, mid, lo;
        split3(vals[2 * kp][cw], vals[2 * kp + 1][cw], hi, mid, lo);
        // 32-bit words: ((blk * 3 + q) * ld2 + c) * 16 + kp
        const int64_t base = (blk * 3 * ld2 + c0 + cw) * 16 + kp;
        out[base] = hi;
        out[base + (int64_t)ld2 * 16] = mid;
        out[base + (int64_t)ld2 * 32] = lo;
    }
}

__global__ __launch_bounds__(256) void k_s_assemble(int P, int ld2, double inv, double eps, const double* __restrict__ G,
                                                    cx<double>* __restrict__ S) {
    // G = O^T O is symmetric and only its lower 128 x 128 tiles are computed:
    // every entry read here has row >= column (G[k][P+l] is read as G[P+l][k]);
    // the upper triangle of S is the conjugate of the lower one.
    const int64_t total = (int64_t)P * P;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(t / P), c = (int)(t % P);
        const int k = r >= c ? r : c, l = r >= c ? c : r;
        const double rr = G[(int64_t)k * ld2 + l], ii = G[(int64_t)(P + k) * ld2 + P + l];
        const double ri = G[(int64_t)(P + l) * ld2 + k], ir = G[(int64_t)(P + k) * ld2 + l];
        const double im = (ri - ir) * inv;
        S[t] = mk((rr + ii) * inv + (k == l ? eps : 0.0), r >= c ? im : -im);
    }
}

template <typename T>
hipError_t launch_o_split3(int64_t s0, int64_t nb, int64_t nbp, int N, int M, int P, int ld2, const void* Trec,
                           const uint64_t* spins, const cx<double>* w, const cx<float>* obar, uint32_t* out,
                           hipStream_t stream) {
    if (nbp % 32 || nb > nbp || ld2 < 2 * P || ld2 % 128 || nbp / 32 > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_o_split3<T>), dim3((unsigned)(ld2 / 128), (unsigned)(nbp / 32)), dim3(256), 0, stream, s0,
                       nb, N, M, P, ld2, Trec, spins, w, obar, out);
    return hipGetLastError();
}

hipError_t launch_s_assemble(int P, int ld2, double inv, double eps, const double* G, cx<double>* S, hipStream_t stream) {
    const int64_t total = (int64_t)P * P;
    const unsigned blocks = (unsigned)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
    hipLaunchKernelGGL(k_s_assemble, dim3(blocks), dim3(256), 0, stream, P, ld2, inv, eps, G, S);
    return hipGetLastError();
}

template hipError_t launch_o_split3<float>(int64_t, int64_t, int64_t, int, int, int, int, const void*, const uint64_t*,
                                           const cx<double>*, const cx<float>*, uint32_t*, hipStream_t);
template hipError_t launch_o_split3<double>(int64_t, int64_t, int64_t, int, int, int, int, const void*, const uint64_t*,
                                            const cx<double>*, const cx<float>*, uint32_t*, hipStream_t);

hipError_t launch_activation(int64_t n, const cx<double>* z, cx<double>* out, int deriv, hipStream_t stream) {
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_activation, dim3((unsigned)g), dim3(256), 0, stream, n, z, out, deriv);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_mc_moves(const RbmDev<T>& md, const McParams& mp, uint64_t* spins, uint32_t* counts, hipStream_t stream) {
    const int jpl = pick_jpl(md.M);
    const dim3 block(256), grid((mp.chains + 15) / 16);
#define RBM_MOVES_CASE(J)                                                                                       \
    case J:                                                                                                     \
        hipLaunchKernelGGL((k_mc_rbm_moves<T, J>), grid, block, 0, stream, md, mp, spins, counts);              \
        break;
    switch (jpl) {
        RBM_JPL_LIST(RBM_MOVES_CASE)
        default: return hipErrorInvalidValue;
    }
#undef RBM_MOVES_CASE
    return hipGetLastError();
}

template hipError_t launch_mc_moves<float>(const RbmDev<float>&, const McParams&, uint64_t*, uint32_t*, hipStream_t);
template hipError_t launch_mc_moves<double>(const RbmDev<double>&, const McParams&, uint64_t*, uint32_t*, hipStream_t);
template hipError_t launch_mc<float>(const RbmDev<float>&, const McParams&, uint64_t*, uint32_t*, bool, size_t, hipStream_t);
template hipError_t launch_mc<double>(const RbmDev<double>&, const McParams&, uint64_t*, uint32_t*, bool, size_t, hipStream_t);
template hipError_t launch_eval<float>(const RbmDev<float>&, const OpDev&, int, int64_t, const uint64_t*, uint64_t, const Records&, bool, size_t, int, hipStream_t);
template hipError_t launch_eval<double>(const RbmDev<double>&, const OpDev&, int, int64_t, const uint64_t*, uint64_t, const Records&, bool, size_t, int, hipStream_t);
template hipError_t launch_ok_reduce<float>(int64_t, int, int, const void*, const uint64_t*, const cx<double>*, int, int64_t, int, cx<double>*, hipStream_t);
template hipError_t launch_ok_reduce<double>(int64_t, int, int, const void*, const uint64_t*, const cx<double>*, int, int64_t, int, cx<double>*, hipStream_t);
template hipError_t launch_s_matrix_ref<float>(int64_t, int, int, int, const void*, const uint64_t*, double, cx<double>*, hipStream_t);
template hipError_t launch_s_matrix_ref<double>(int64_t, int, int, int, const void*, const uint64_t*, double, cx<double>*, hipStream_t);

}  // namespace rbmhip
