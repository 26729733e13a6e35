// Host side of librbmhip.so: the C ABI declared in include/rbmhip.h.
//
// Objects (wavefunctions, operators, ensembles) are host descriptors; device
// copies are made lazily per device and refreshed after setters.  Each call
// runs its kernels on one stream of the current device, reduces in FP64 on
// the device and synchronises before returning (the reference's calls block on
// a device-to-host copy, include/Array.hpp / source/Array.cu:96-100).
#include "../../include/rbmhip.h"

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "rbm_kernels.hpp"
#include "deep_kernels.hpp"
#include "made_kernels.hpp"
#include "linalg_kernels.hpp"
#include "noalloc.hpp"

using cd = std::complex<double>;

namespace rbmhip {

thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define RBM_HIPCHK(x)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess)                                                                          \
            throw ::rbmhip::Error(RBMHIP_ERR_HIP, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

#define RBM_REQUIRE(cond, msg)                                              \
    do {                                                                    \
        if (!(cond)) throw ::rbmhip::Error(RBMHIP_ERR_ARG, std::string(msg)); \
    } while (0)

// A roctx range around a compute entry point, so that a profiler's trace
// (rocprofv3 --marker-trace) groups the kernels of each call under its name.
struct ApiRange {
    explicit ApiRange(const char* name) { roctxRangePush(name); }
    ~ApiRange() { roctxRangePop(); }
    ApiRange(const ApiRange&) = delete;
    ApiRange& operator=(const ApiRange&) = delete;
};

template <class F>
int guard(F&& f) {
    try {
        f();
        return RBMHIP_OK;
    } catch (Error& e) {
        g_err = e.what();
        return e.code;
    } catch (std::bad_alloc&) {
        g_err = "host allocation failed";
        return RBMHIP_ERR_ARG;
    } catch (std::exception& e) {
        g_err = e.what();
        return RBMHIP_ERR_ARG;
    }
}

// ---------------------------------------------------------------------------
// device memory and per-device context
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    void* ensure(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes > n) {
            if (p) RBM_HIPCHK(hipFree(p));
            p = nullptr;
            n = 0;
            RBM_HIPCHK(hipMalloc(&p, bytes));
            n = bytes;
        }
        return p;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t max_lds = 65536;
    int num_cus = 256;
    // device address of this device's allocation-free scratch array (null
    // with scratch_err set when the symbol could not be resolved)
    void* scratch = nullptr;
    std::string scratch_err;
    std::map<std::string, DevBuf> ws;
    bool timing = false;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::map<std::string, std::pair<double, int64_t>> totals;
    // second stream (default priority) for work that overlaps the main one,
    // with two events that order the streams; created on first use
    hipStream_t side = nullptr;
    hipEvent_t ev_main = nullptr, ev_side = nullptr;
    void ensure_side() {
        if (side) return;
        RBM_HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        RBM_HIPCHK(hipEventCreateWithFlags(&ev_main, hipEventDisableTiming));
        RBM_HIPCHK(hipEventCreateWithFlags(&ev_side, hipEventDisableTiming));
    }

    void* buf(const char* name, size_t bytes) { return ws[name].ensure(bytes); }

    template <class F>
    void launch(const char* name, F&& f) {
        if (!timing) {
            RBM_HIPCHK(f());
            return;
        }
        Pending p{name, nullptr, nullptr};
        RBM_HIPCHK(hipEventCreate(&p.a));
        RBM_HIPCHK(hipEventCreate(&p.b));
        RBM_HIPCHK(hipEventRecord(p.a, stream));
        RBM_HIPCHK(f());
        RBM_HIPCHK(hipEventRecord(p.b, stream));
        pending.push_back(p);
    }

    void sync() {
        RBM_HIPCHK(hipStreamSynchronize(stream));
        for (auto& p : pending) {
            float ms = 0.f;
            RBM_HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
            auto& t = totals[p.name];
            t.first += ms;
            t.second += 1;
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        pending.clear();
    }

    void d2h(void* dst, const void* src, size_t bytes) {
        RBM_HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    }
    void h2d(void* dst, const void* src, size_t bytes) {
        RBM_HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    }
};

static std::mutex g_mu;
static int g_device = 0;
static std::map<int, std::unique_ptr<Ctx>> g_ctx;

static Ctx& ctx() {
    std::lock_guard<std::mutex> lk(g_mu);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        throw Error(RBMHIP_ERR_NO_DEVICE, "no HIP device is visible; librbmhip runs on the GPU only");
    RBM_REQUIRE(g_device >= 0 && g_device < count, "device index out of range");
    RBM_HIPCHK(hipSetDevice(g_device));
    auto it = g_ctx.find(g_device);
    if (it != g_ctx.end()) return *it->second;
    auto c = std::make_unique<Ctx>();
    c->device = g_device;
    RBM_HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, g_device) == hipSuccess && v > 0)
        c->max_lds = (size_t)v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, g_device) == hipSuccess && v > 0)
        c->num_cus = v;
    const hipError_t se = noalloc_scratch_address(&c->scratch);
    if (se != hipSuccess) {
        c->scratch = nullptr;
        c->scratch_err = hipGetErrorString(se);
    }
    Ctx& ref = *c;
    g_ctx[g_device] = std::move(c);
    return ref;
}

}  // namespace rbmhip

using namespace rbmhip;

// ---------------------------------------------------------------------------
// wavefunction descriptors
// ---------------------------------------------------------------------------
struct DevModel {
    uint64_t version = ~(uint64_t)0;
    int precision = 0;
    DevBuf W, b, a;
    // PsiDeep
    DevBuf conns, weights, biases, final_w, rhs_begin, rhs_entries;
    // PsiDeep, BF16 MFMA layers
    uint64_t mfma_version = ~(uint64_t)0;
    DevBuf mfrag, mvec;
};

struct rbmhip_psi {
    int kind = 0;  // 0 RBM, 1 PsiDeep
    int N = 0, M = 0;
    int precision = RBMHIP_FP32;
    double prefactor = 1.0;
    std::vector<cd> b, W, a;
    bool has_a = false;
    // PsiDeep
    std::vector<int32_t> sizes, conn_counts;
    std::vector<uint32_t> conns;
    std::vector<cd> weights, biases, final_weights;
    std::vector<int> rhs_begin, rhs_entries, rb_off;
    int translational_invariance = 0;
    int ref_compat_final = 1;
    uint64_t version = 0;
    std::map<int, DevModel> dev;

    int64_t num_params() const {
        if (kind == 0) return 2 * (int64_t)N + M + (int64_t)N * M + (has_a ? N : 0);
        int64_t p = 2 * (int64_t)N;
        for (size_t l = 0; l < sizes.size(); l++) p += sizes[l] + (int64_t)conn_counts[l] * sizes[l];
        return p + sizes.back();
    }
    ~rbmhip_psi() {
        for (auto& kv : dev) {
            kv.second.W.release(); kv.second.b.release(); kv.second.a.release();
            kv.second.conns.release(); kv.second.weights.release(); kv.second.biases.release();
            kv.second.final_w.release(); kv.second.rhs_begin.release(); kv.second.rhs_entries.release();
            kv.second.mfrag.release(); kv.second.mvec.release();
        }
    }
};

template <typename T>
static void upload_cx(Ctx& c, DevBuf& buf, const std::vector<cd>& v) {
    std::vector<cx<T>> h(v.size());
    for (size_t i = 0; i < v.size(); i++) h[i] = mk<T>((T)v[i].real(), (T)v[i].imag());
    buf.ensure(sizeof(cx<T>) * h.size());
    if (!h.empty()) RBM_HIPCHK(hipMemcpy(buf.p, h.data(), sizeof(cx<T>) * h.size(), hipMemcpyHostToDevice));
    (void)c;
}

template <typename T>
static RbmDev<T> rbm_dev(Ctx& c, rbmhip_psi* psi) {
    DevModel& d = psi->dev[c.device];
    const int ld = rbm_padded_width(psi->M);
    if (d.version != psi->version || d.precision != psi->precision) {
        // W and b padded with zero columns to ld hidden units
        std::vector<cd> Wp((size_t)psi->N * ld, cd(0, 0)), bp((size_t)ld, cd(0, 0));
        for (int i = 0; i < psi->N; i++)
            for (int j = 0; j < psi->M; j++) Wp[(size_t)i * ld + j] = psi->W[(size_t)i * psi->M + j];
        for (int j = 0; j < psi->M; j++) bp[j] = psi->b[j];
        upload_cx<T>(c, d.W, Wp);
        upload_cx<T>(c, d.b, bp);
        if (psi->has_a) upload_cx<T>(c, d.a, psi->a);
        d.version = psi->version;
        d.precision = psi->precision;
    }
    RbmDev<T> md;
    md.N = psi->N;
    md.M = psi->M;
    md.ld = ld;
    md.W = reinterpret_cast<const cx<T>*>(d.W.p);
    md.b = reinterpret_cast<const cx<T>*>(d.b.p);
    md.a = psi->has_a ? reinterpret_cast<const cx<T>*>(d.a.p) : nullptr;
    return md;
}

// bfloat16 bits of x, round to nearest even (finite inputs)
static uint16_t bf16_bits(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static float bf16_round(double x) {
    const uint32_t u = (uint32_t)bf16_bits((float)x) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

template <typename T>
static DeepDev<T> deep_dev(Ctx& c, rbmhip_psi* psi) {
    DevModel& d = psi->dev[c.device];
    const int L = (int)psi->sizes.size();
    if (d.version != psi->version || d.precision != psi->precision) {
        if (psi->precision == RBMHIP_BF16) {
            // the complex64 kernels (O_k) see the same bfloat16 weights as the MFMA layers
            std::vector<cd> wr(psi->weights.size());
            for (size_t i = 0; i < wr.size(); i++)
                wr[i] = cd(bf16_round(psi->weights[i].real()), bf16_round(psi->weights[i].imag()));
            upload_cx<T>(c, d.weights, wr);
        } else {
            upload_cx<T>(c, d.weights, psi->weights);
        }
        upload_cx<T>(c, d.biases, psi->biases);
        upload_cx<T>(c, d.final_w, psi->final_weights);
        d.conns.ensure(sizeof(uint32_t) * psi->conns.size());
        RBM_HIPCHK(hipMemcpy(d.conns.p, psi->conns.data(), sizeof(uint32_t) * psi->conns.size(), hipMemcpyHostToDevice));
        d.rhs_begin.ensure(sizeof(int) * psi->rhs_begin.size());
        if (!psi->rhs_begin.empty())
            RBM_HIPCHK(hipMemcpy(d.rhs_begin.p, psi->rhs_begin.data(), sizeof(int) * psi->rhs_begin.size(), hipMemcpyHostToDevice));
        d.rhs_entries.ensure(sizeof(int) * psi->rhs_entries.size());
        if (!psi->rhs_entries.empty())
            RBM_HIPCHK(hipMemcpy(d.rhs_entries.p, psi->rhs_entries.data(), sizeof(int) * psi->rhs_entries.size(), hipMemcpyHostToDevice));
        d.version = psi->version;
        d.precision = psi->precision;
    }
    DeepDev<T> md;
    md.N = psi->N;
    md.num_layers = L;
    int64_t woff = 0, boff = 0;
    int width = psi->N, units = 0;
    for (int l = 0; l < L; l++) {
        md.size[l] = psi->sizes[l];
        md.conn[l] = psi->conn_counts[l];
        md.w_off[l] = woff;
        md.b_off[l] = boff;
        woff += (int64_t)psi->conn_counts[l] * psi->sizes[l];
        boff += psi->sizes[l];
        width = std::max(width, (int)psi->sizes[l]);
        units += psi->sizes[l];
        md.rb_off[l] = l < (int)psi->rb_off.size() ? psi->rb_off[l] : 0;
    }
    md.rhs_begin = reinterpret_cast<const int*>(d.rhs_begin.p);
    md.rhs_entries = reinterpret_cast<const int*>(d.rhs_entries.p);
    md.width = width;
    md.units = units;
    md.conns = reinterpret_cast<const uint32_t*>(d.conns.p);
    md.weights = reinterpret_cast<const cx<T>*>(d.weights.p);
    md.biases = reinterpret_cast<const cx<T>*>(d.biases.p);
    md.final_w = reinterpret_cast<const cx<T>*>(d.final_w.p);
    md.translational_invariance = psi->translational_invariance;
    md.ref_compat_final = psi->ref_compat_final;
    return md;
}

// Dense BF16 fragments of a PsiDeep's layers (layout in deep_kernels.hpp).
static bool deep_mfma_fits(const rbmhip_psi* psi, std::string* why) {
    if (psi->translational_invariance) { if (why) *why = "translational invariance"; return false; }
    int prev = psi->N;
    for (size_t l = 0; l < psi->sizes.size(); l++) {
        if ((prev + 31) / 32 > DM_MAX_KB || (psi->sizes[l] + 31) / 32 > DM_MAX_KB) {
            if (why) *why = "a layer wider than 256";
            return false;
        }
        prev = psi->sizes[l];
    }
    return true;
}

static DeepMfmaDev deep_mfma_dev(Ctx& c, rbmhip_psi* psi) {
    std::string why;
    if (!deep_mfma_fits(psi, &why))
        throw Error(RBMHIP_ERR_UNSUPPORTED, "precision=16 (BF16 MFMA) PsiDeep does not support " + why);
    const int L = (int)psi->sizes.size();
    DeepMfmaDev md;
    std::memset(&md, 0, sizeof(md));
    md.N = psi->N;
    md.L = L;
    int fo = 0, vo = 0, prev = psi->N;
    for (int l = 0; l < L; l++) {
        md.in_p[l] = (prev + 31) / 32 * 32;
        md.out_p[l] = l == L - 1 ? (psi->sizes[l] + 15) / 16 * 16 : (psi->sizes[l] + 31) / 32 * 32;
        md.frag_off[l] = fo;
        md.vec_off[l] = vo;
        fo += (md.out_p[l] / 16) * (md.in_p[l] / 32) * 2 * 64;
        vo += 2 * md.out_p[l];
        prev = psi->sizes[l];
    }
    md.final_off = vo;
    vo += 2 * md.out_p[L - 1];
    md.fwd_total = fo;
    // backward fragments: W_{l+1} itself, units of layer l x units of layer l+1
    for (int l = 0; l + 1 < L; l++) {
        md.bfrag_off[l] = fo;
        fo += (md.out_p[l] / 16) * ((md.out_p[l + 1] + 31) / 32) * 2 * 64;
    }
    md.frag_total = fo;
    md.vec_total = vo;
    int ro = 0;
    for (int l = 0; l < L; l++) { md.rec_x[l] = ro; ro += md.out_p[l]; }
    for (int l = 0; l < L; l++) { md.rec_d[l] = ro; ro += md.out_p[l]; }
    md.rec_lp = ro;
    md.rec_stride = ro + 1;
    RBM_REQUIRE(deep_mfma_lds_bytes(md) <= c.max_lds,
                "the bfloat16 layers need " + std::to_string(deep_mfma_lds_bytes(md)) + " bytes of LDS");
    DevModel& d = psi->dev[c.device];
    if (d.mfma_version != psi->version || d.mfrag.p == nullptr) {
        std::vector<uint16_t> frag((size_t)fo * 8, 0);
        std::vector<float> vec((size_t)vo, 0.0f);
        int64_t woff = 0, boff = 0;
        prev = psi->N;
        std::vector<std::vector<cd>> dense((size_t)L);
        for (int l = 0; l < L; l++) {
            const int size = psi->sizes[l], C = psi->conn_counts[l];
            // dense W[i][j] (inputs x units) from the connection table
            std::vector<cd>& Wd = dense[(size_t)l];
            Wd.assign((size_t)prev * size, cd(0, 0));
            for (int cc = 0; cc < C; cc++)
                for (int j = 0; j < size; j++) {
                    const int64_t slot = woff + (int64_t)cc * size + j;
                    Wd[(size_t)psi->conns[slot] * size + j] += psi->weights[slot];
                }
            const int KB = md.in_p[l] / 32, MT = md.out_p[l] / 16;
            for (int mt = 0; mt < MT; mt++)
                for (int kb = 0; kb < KB; kb++)
                    for (int part = 0; part < 2; part++)
                        for (int lane = 0; lane < 64; lane++)
                            for (int v = 0; v < 8; v++) {
                                const int m = mt * 16 + (lane & 15), g = lane >> 4;
                                const int k = l == 0 ? kb * 32 + 8 * g + v
                                                     : kb * 32 + (v < 4 ? 4 * g + v : 16 + 4 * g + (v - 4));
                                float w = 0.0f;
                                if (m < size && k < prev) {
                                    const cd z = Wd[(size_t)k * size + m];
                                    w = (float)(part == 0 ? z.real() : z.imag());
                                }
                                const size_t at = ((size_t)md.frag_off[l] + (size_t)((mt * KB + kb) * 2 + part) * 64 + lane) * 8 + v;
                                frag[at] = bf16_bits(w);
                            }
            for (int j = 0; j < size; j++) {
                vec[(size_t)md.vec_off[l] + j] = (float)psi->biases[boff + j].real();
                vec[(size_t)md.vec_off[l] + md.out_p[l] + j] = (float)psi->biases[boff + j].imag();
            }
            woff += (int64_t)C * size;
            boff += size;
            prev = size;
        }
        for (int j = 0; j < psi->sizes[L - 1]; j++) {
            vec[(size_t)md.final_off + j] = (float)psi->final_weights[j].real();
            vec[(size_t)md.final_off + md.out_p[L - 1] + j] = (float)psi->final_weights[j].imag();
        }
        for (int l = 0; l + 1 < L; l++) {
            // A[m][k] = W_{l+1}[m][k]: m a unit of layer l, k a unit of layer l+1
            const std::vector<cd>& Wd = dense[(size_t)l + 1];
            const int rows = psi->sizes[l], cols = psi->sizes[l + 1];
            const int MT = md.out_p[l] / 16, KB = (md.out_p[l + 1] + 31) / 32;
            for (int mt = 0; mt < MT; mt++)
                for (int kb = 0; kb < KB; kb++)
                    for (int part = 0; part < 2; part++)
                        for (int lane = 0; lane < 64; lane++)
                            for (int v = 0; v < 8; v++) {
                                const int m = mt * 16 + (lane & 15), g = lane >> 4;
                                const int k = kb * 32 + (v < 4 ? 4 * g + v : 16 + 4 * g + (v - 4));
                                float w = 0.0f;
                                if (m < rows && k < cols) {
                                    const cd z = Wd[(size_t)m * cols + k];
                                    w = (float)(part == 0 ? z.real() : z.imag());
                                }
                                const size_t at = ((size_t)md.bfrag_off[l] + (size_t)((mt * KB + kb) * 2 + part) * 64 + lane) * 8 + v;
                                frag[at] = bf16_bits(w);
                            }
        }
        d.mfrag.ensure(frag.size() * sizeof(uint16_t));
        RBM_HIPCHK(hipMemcpy(d.mfrag.p, frag.data(), frag.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        d.mvec.ensure(vec.size() * sizeof(float));
        RBM_HIPCHK(hipMemcpy(d.mvec.p, vec.data(), vec.size() * sizeof(float), hipMemcpyHostToDevice));
        d.mfma_version = psi->version;
    }
    md.frag = reinterpret_cast<const uint4*>(d.mfrag.p);
    md.vec = reinterpret_cast<const float*>(d.mvec.p);
    return md;
}

// ---------------------------------------------------------------------------
// operators: canonical flip-mask form of the Pauli table
// ---------------------------------------------------------------------------
struct DevOp {
    DevBuf diag_coef, diag_zy, flip_begin, flips, ent_begin, ent_coef, ent_zy;
    bool ready = false;
};

struct rbmhip_op {
    int S = 0, L = 0;
    std::vector<cd> coeff;
    std::vector<int32_t> types, idx;
    // canonical tables
    std::vector<cx<double>> diag_coef;
    std::vector<uint64_t> diag_zy;
    std::vector<int> flip_begin{0}, flips, ent_begin{0};
    std::vector<cx<double>> ent_coef;
    std::vector<uint64_t> ent_zy;
    int max_index = -1;
    bool hermitian = true;
    std::map<int, DevOp> dev;
    ~rbmhip_op() {
        for (auto& kv : dev) {
            DevOp& d = kv.second;
            d.diag_coef.release(); d.diag_zy.release(); d.flip_begin.release(); d.flips.release();
            d.ent_begin.release(); d.ent_coef.release(); d.ent_zy.release();
        }
    }
};

static void bit_set(uint64_t* w, int i) { w[i >> 6] ^= (uint64_t)1 << (i & 63); }

// Canonicalisation of include/operator/Operator.hpp:45-88: X and Y flip the
// spin, Y multiplies by -i s, Z by s, with s the ORIGINAL configuration.  The
// string's element is coef (-i)^{#Y} prod_{i in zy} s_i with zy the XOR of
// the Y and Z positions (s_i^2 = 1), and its target is s ^ (XOR of the X and
// Y positions).
static void canonicalise(rbmhip_op* op) {
    struct Ent { uint64_t mask[2]; uint64_t zy[2]; cd c; };
    std::vector<Ent> ents;
    const int64_t table = (int64_t)op->S * op->L;
    for (int n = 0; n < op->S; n++) {
        Ent e{{0, 0}, {0, 0}, op->coeff[n]};
        int ny = 0;
        int64_t t = (int64_t)n * op->L;
        for (;; t++) {
            RBM_REQUIRE(t < table, "operator string " + std::to_string(n) + " has no -1 terminator before the end of the table");
            const int pidx = op->idx[t];
            if (pidx == -1) break;
            const int pt = op->types[t];
            if (pt == 0) continue;
            RBM_REQUIRE(pidx >= 0 && pidx < 128, "Pauli index out of range [0, 128)");
            op->max_index = std::max(op->max_index, pidx);
            if (pt == 1) { bit_set(e.mask, pidx); }
            else if (pt == 2) { bit_set(e.mask, pidx); bit_set(e.zy, pidx); ny++; }
            else if (pt == 3) { bit_set(e.zy, pidx); }
        }
        static const cd minus_i_pow[4] = {cd(1, 0), cd(0, -1), cd(-1, 0), cd(0, 1)};
        e.c *= minus_i_pow[ny & 3];
        ents.push_back(e);
    }
    // group by flip mask (first-appearance order), diagonal separately
    std::vector<std::pair<uint64_t, uint64_t>> masks;
    std::vector<std::vector<Ent>> groups;
    for (const Ent& e : ents) {
        if (e.mask[0] == 0 && e.mask[1] == 0) {
            op->diag_coef.push_back(mk(e.c.real(), e.c.imag()));
            op->diag_zy.push_back(e.zy[0]);
            op->diag_zy.push_back(e.zy[1]);
            continue;
        }
        size_t g = 0;
        for (; g < masks.size(); g++)
            if (masks[g].first == e.mask[0] && masks[g].second == e.mask[1]) break;
        if (g == masks.size()) { masks.push_back({e.mask[0], e.mask[1]}); groups.emplace_back(); }
        groups[g].push_back(e);
    }
    for (size_t g = 0; g < masks.size(); g++) {
        for (int i = 0; i < 128; i++)
            if (((i < 64 ? masks[g].first >> i : masks[g].second >> (i - 64)) & 1u)) op->flips.push_back(i);
        op->flip_begin.push_back((int)op->flips.size());
        for (const Ent& e : groups[g]) {
            op->ent_coef.push_back(mk(e.c.real(), e.c.imag()));
            op->ent_zy.push_back(e.zy[0]);
            op->ent_zy.push_back(e.zy[1]);
        }
        op->ent_begin.push_back((int)op->ent_coef.size());
    }
    // Hermiticity: within a flip mask m the element at s is sum_z C_z chi_z(s)
    // with chi_z(s) = prod_{i in z} s_i; <s|H|s^m> = conj(<s^m|H|s>) for all s
    // iff C_z = (-1)^{|z & m|} conj(C_z) for every zy pattern z.
    auto check = [&](uint64_t m0, uint64_t m1, const std::vector<Ent>& es) {
        std::vector<std::pair<std::pair<uint64_t, uint64_t>, cd>> acc;
        double scale = 0.0;
        for (const Ent& e : es) {
            scale = std::max(scale, std::abs(e.c));
            bool found = false;
            for (auto& a : acc)
                if (a.first.first == e.zy[0] && a.first.second == e.zy[1]) { a.second += e.c; found = true; break; }
            if (!found) acc.push_back({{e.zy[0], e.zy[1]}, e.c});
        }
        for (auto& a : acc) {
            const int par = (__builtin_popcountll(a.first.first & m0) + __builtin_popcountll(a.first.second & m1)) & 1;
            const cd target = (par ? -1.0 : 1.0) * std::conj(a.second);
            if (std::abs(a.second - target) > 1e-12 * std::max(1.0, scale)) op->hermitian = false;
        }
    };
    std::vector<Ent> diag;
    for (const Ent& e : ents)
        if (e.mask[0] == 0 && e.mask[1] == 0) diag.push_back(e);
    check(0, 0, diag);
    for (size_t g = 0; g < masks.size(); g++) check(masks[g].first, masks[g].second, groups[g]);
}

static OpDev op_dev(Ctx& c, rbmhip_op* op) {
    DevOp& d = op->dev[c.device];
    auto up = [&](DevBuf& b, const void* src, size_t bytes) {
        b.ensure(bytes);
        if (bytes) RBM_HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    if (!d.ready) {
        up(d.diag_coef, op->diag_coef.data(), op->diag_coef.size() * sizeof(cx<double>));
        up(d.diag_zy, op->diag_zy.data(), op->diag_zy.size() * sizeof(uint64_t));
        up(d.flip_begin, op->flip_begin.data(), op->flip_begin.size() * sizeof(int));
        up(d.flips, op->flips.data(), op->flips.size() * sizeof(int));
        up(d.ent_begin, op->ent_begin.data(), op->ent_begin.size() * sizeof(int));
        up(d.ent_coef, op->ent_coef.data(), op->ent_coef.size() * sizeof(cx<double>));
        up(d.ent_zy, op->ent_zy.data(), op->ent_zy.size() * sizeof(uint64_t));
        d.ready = true;
    }
    OpDev o;
    o.n_diag = (int)op->diag_coef.size();
    o.diag_coef = reinterpret_cast<const cx<double>*>(d.diag_coef.p);
    o.diag_zy = reinterpret_cast<const uint64_t*>(d.diag_zy.p);
    o.n_terms = (int)op->flip_begin.size() - 1;
    o.flip_begin = reinterpret_cast<const int*>(d.flip_begin.p);
    o.flips = reinterpret_cast<const int*>(d.flips.p);
    o.ent_begin = reinterpret_cast<const int*>(d.ent_begin.p);
    o.ent_coef = reinterpret_cast<const cx<double>*>(d.ent_coef.p);
    o.ent_zy = reinterpret_cast<const uint64_t*>(d.ent_zy.p);
    return o;
}

static OpDev empty_op() {
    OpDev o;
    std::memset(&o, 0, sizeof(o));
    return o;
}

// ---------------------------------------------------------------------------
// masked autoregressive proposal (MADE)
// ---------------------------------------------------------------------------
// Device parameters are one padded FP32 block [W1 (Hp x Np), b1 (Hp), W2 (Np x Hp),
// b2 (Np)] so that Adam runs over a single array; masked and padded entries
// are zero and receive zero gradients, so they stay zero.
struct rbmhip_made {
    MadeShape s{};
    int device = -1;
    int64_t off_b1 = 0, off_W2 = 0, off_b2 = 0, padded = 0;
    DevBuf params, grad, m, v;
    int t = 0;                  // Adam steps taken
    DevBuf W1t_bf, W2_bf, b1_sorted;   // bfloat16 inference copies, sampler layout
    bool inf_valid = false;
    ~rbmhip_made() {
        params.release(); grad.release(); m.release(); v.release();
        W1t_bf.release(); W2_bf.release(); b1_sorted.release();
    }
    int64_t num_params() const { return 2 * (int64_t)s.N * s.H + s.H + s.N; }
    float* P() { return reinterpret_cast<float*>(params.p); }
    float* G() { return reinterpret_cast<float*>(grad.p); }
};

static void made_check_device(Ctx& c, const rbmhip_made* md) {
    RBM_REQUIRE(c.device == md->device, "the proposal network lives on another device");
}

// host flat layout [W1 (H x N), b1, W2 (N x H), b2] <-> padded device layout, masks applied
static std::vector<float> made_pad_params(const rbmhip_made* md, const float* flat) {
    const int N = md->s.N, H = md->s.H, Np = md->s.Np, Hp = md->s.Hp;
    std::vector<float> out((size_t)md->padded, 0.0f);
    const float* W1 = flat;
    const float* b1 = W1 + (size_t)H * N;
    const float* W2 = b1 + H;
    const float* b2 = W2 + (size_t)N * H;
    for (int k = 0; k < H; k++) {
        const int deg = made_degree(k, N);
        for (int i = 0; i < N; i++) out[(size_t)k * Np + i] = i < deg ? W1[(size_t)k * N + i] : 0.0f;
        out[md->off_b1 + k] = b1[k];
    }
    for (int j = 0; j < N; j++) {
        for (int k = 0; k < H; k++) out[md->off_W2 + (size_t)j * Hp + k] = made_degree(k, N) <= j ? W2[(size_t)j * H + k] : 0.0f;
        out[md->off_b2 + j] = b2[j];
    }
    return out;
}

static std::vector<float> made_unpad_params(const rbmhip_made* md, const std::vector<float>& dev) {
    const int N = md->s.N, H = md->s.H, Np = md->s.Np, Hp = md->s.Hp;
    std::vector<float> out((size_t)md->num_params());
    float* W1 = out.data();
    float* b1 = W1 + (size_t)H * N;
    float* W2 = b1 + H;
    float* b2 = W2 + (size_t)N * H;
    for (int k = 0; k < H; k++) {
        for (int i = 0; i < N; i++) W1[(size_t)k * N + i] = dev[(size_t)k * Np + i];
        b1[k] = dev[md->off_b1 + k];
    }
    for (int j = 0; j < N; j++) {
        for (int k = 0; k < H; k++) W2[(size_t)j * H + k] = dev[md->off_W2 + (size_t)j * Hp + k];
        b2[j] = dev[md->off_b2 + j];
    }
    return out;
}

// Seeded initialisation: parameter k (flat host index) draws Philox(seed; ctr =
// (k, 0, 0, 3)) word x, u = (x + 0.5) 2^-32; weights (2u - 1) / sqrt(fan_in)
// (fan_in N for W1, H for W2), biases zero; masks applied afterwards.
static std::vector<float> made_init_params(int N, int H, uint64_t seed) {
    const int64_t P = 2 * (int64_t)N * H + H + N;
    std::vector<float> out((size_t)P, 0.0f);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const int64_t w1 = (int64_t)H * N, w2_begin = w1 + H, w2_end = w2_begin + (int64_t)N * H;
    for (int64_t k = 0; k < P; k++) {
        const bool is_w1 = k < w1, is_w2 = k >= w2_begin && k < w2_end;
        if (!is_w1 && !is_w2) continue;
        const u32x4 r = philox((uint32_t)k, 0u, 0u, 3u, k0, k1);
        const double u = ((double)r.x + 0.5) * (1.0 / 4294967296.0);
        const double fan = is_w1 ? (double)N : (double)H;
        out[(size_t)k] = (float)((2.0 * u - 1.0) / std::sqrt(fan));
    }
    return out;
}

static MadeInf made_inf(Ctx& c, rbmhip_made* md) {
    made_check_device(c, md);
    const size_t bytes = (size_t)md->s.N * md->s.Hp * sizeof(uint16_t);
    md->W1t_bf.ensure(bytes);
    md->W2_bf.ensure(bytes);
    md->b1_sorted.ensure((size_t)md->s.Hp * sizeof(float));
    if (!md->inf_valid) {
        const float* P = md->P();
        c.launch("made_inf_weights", [&] {
            return launch_made_inference_weights(md->s, P, P + md->off_b1, P + md->off_W2,
                                                 reinterpret_cast<uint16_t*>(md->W1t_bf.p),
                                                 reinterpret_cast<uint16_t*>(md->W2_bf.p),
                                                 reinterpret_cast<float*>(md->b1_sorted.p), c.stream);
        });
        md->inf_valid = true;
    }
    MadeInf mi;
    mi.N = md->s.N; mi.H = md->s.H; mi.Hp = md->s.Hp;
    mi.W1t = reinterpret_cast<const uint16_t*>(md->W1t_bf.p);
    mi.W2 = reinterpret_cast<const uint16_t*>(md->W2_bf.p);
    mi.b1 = reinterpret_cast<const float*>(md->b1_sorted.p);
    mi.b2 = md->P() + md->off_b2;
    return mi;
}

// Forward and backward pass of loss = -(1/B) sum_s log q(x_s) over B rows of X
// (Bp = padded rows).  Leaves the gradient in md->grad and the per-tile log-q
// partials in `partials`.
// Launch extents of one training step (gather, forward and backward GEMMs,
// bias column sums, Adam): for every buffer a kernel of the step touches, the
// largest element offset its launch arguments allow, against the size the
// step allocates for that buffer.  Pure host arithmetic, run before every
// training call and by rbmhip_made_check_launches without a GPU.
static int made_splits(int Bp) {
    int splits = 1;
    while (splits < 64 && Bp % (16 * splits * 2) == 0 && Bp / (splits * 2) >= 64) splits *= 2;
    return splits;
}

struct MadeStepPlan {
    int N, H, Np, Hp, B, Bp, splits, tiles;
    int64_t n, off_b1, off_W2, off_b2, padded;
};

static MadeStepPlan made_plan(int N, int H, int64_t n, int batch) {
    MadeStepPlan p{};
    p.N = N; p.H = H; p.Np = made_pad(N); p.Hp = made_pad(H);
    p.B = batch; p.Bp = made_pad(batch); p.n = n;
    p.splits = made_splits(p.Bp);
    p.tiles = (p.Np / 64) * (p.Bp / 64);
    p.off_b1 = (int64_t)p.Hp * p.Np;
    p.off_W2 = p.off_b1 + p.Hp;
    p.off_b2 = p.off_W2 + (int64_t)p.Np * p.Hp;
    p.padded = p.off_b2 + p.Np;
    return p;
}

static void made_check_extents(const MadeStepPlan& p) {
    auto need = [](int64_t touched, int64_t size, const char* what) {
        if (touched > size)
            throw Error(RBMHIP_ERR_ARG, std::string("launch extent exceeds its buffer: ") + what + " touches " +
                                            std::to_string(touched) + " elements of " + std::to_string(size));
    };
    // k_gemm<EPI> over an M x N output, K deep, split `splits` ways: A at
    // (m sam + k sak), B at (k sbk + n sbn), C at (z M ldc + m ldc + n)
    auto gemm = [&](const char* what, int M, int N, int K, int64_t sam, int64_t sak, int64_t a_size, int64_t sbk,
                    int64_t sbn, int64_t b_size, int64_t ldc, int64_t c_size, int splits) {
        RBM_REQUIRE(M % 64 == 0 && N % 64 == 0 && K % (16 * splits) == 0, std::string("GEMM shape: ") + what);
        need((M - 1) * sam + (K - 1) * sak + 1, a_size, what);
        need((K - 1) * sbk + (N - 1) * sbn + 1, b_size, what);
        need((int64_t)(splits - 1) * M * ldc + (M - 1) * ldc + N, c_size, what);
    };
    const int64_t X = (int64_t)p.Bp * p.Np, HB = (int64_t)p.Bp * p.Hp, W = (int64_t)p.Np * p.Hp;
    // minibatch indices and the gather of Bp rows of Np inputs
    need(p.B, p.B, "batch indices");
    need(X, X, "gathered inputs");
    need(2 * p.n, 2 * p.n, "spin words");
    // a = X W1^T + b1 (W1: Hp x Np at 0, b1 at off_b1), h = relu(a)
    gemm("a = X W1^T", p.Bp, p.Hp, p.Np, p.Np, 1, X, 1, p.Np, p.off_b1, p.Hp, HB, 1);
    need(p.off_b1 + p.Hp, p.padded, "b1");
    // l = h W2^T + b2 (W2: Np x Hp at off_W2), loss partials per tile
    gemm("l = h W2^T", p.Bp, p.Np, p.Hp, p.Hp, 1, HB, 1, p.Hp, W, p.Np, X, 1);
    need(p.off_W2 + W, p.padded, "W2");
    need(p.off_b2 + p.Np, p.padded, "b2");
    need((p.Bp / 64) * (p.Np / 64), p.tiles, "loss partials");
    // dW2 = dl^T h split over K, reduced into G at off_W2
    gemm("dW2 = dl^T h", p.Np, p.Hp, p.Bp, 1, p.Np, X, p.Hp, 1, HB, p.Hp, (int64_t)p.splits * W, p.splits);
    // dh = (dl W2) relu'(a)
    gemm("dh = dl W2", p.Bp, p.Hp, p.Np, p.Np, 1, X, p.Hp, 1, W, p.Hp, HB, 1);
    // dW1 = dh^T X split over K, reduced into G at 0
    gemm("dW1 = dh^T X", p.Hp, p.Np, p.Bp, 1, p.Hp, HB, p.Np, 1, X, p.Np, (int64_t)p.splits * W, p.splits);
    // bias gradients: column sums of dh (Hp columns) and dl (Np columns)
    need((int64_t)COLSUM_SLICES * std::max(p.Hp, p.Np), (int64_t)COLSUM_SLICES * std::max(p.Hp, p.Np), "column-sum slices");
    need(p.off_b1 + p.Hp, p.padded, "db1");
    need(p.off_b2 + p.Np, p.padded, "db2");
}

static void made_forward_backward(Ctx& c, rbmhip_made* md, int B, int Bp, const float* X, double* partials) {
    const int N = md->s.N, H = md->s.H, Np = md->s.Np, Hp = md->s.Hp;
    float* Apre = reinterpret_cast<float*>(c.buf("made_apre", (size_t)Bp * Hp * sizeof(float)));
    float* Hh = reinterpret_cast<float*>(c.buf("made_h", (size_t)Bp * Hp * sizeof(float)));
    float* dL = reinterpret_cast<float*>(c.buf("made_dl", (size_t)Bp * Np * sizeof(float)));
    float* dH = reinterpret_cast<float*>(c.buf("made_dh", (size_t)Bp * Hp * sizeof(float)));
    const float* P = md->P();
    float* G = md->G();
    const float* W1 = P;
    const float* W2 = P + md->off_W2;
    GemmEpiArgs e{};
    // a = X W1^T + b1, h = relu(a)
    e.bias = P + md->off_b1; e.C2 = Hh;
    c.launch("made_gemm", [&] { return launch_gemm(EPI_BIAS_RELU, Bp, Hp, Np, X, Np, 1, W1, 1, Np, Apre, Hp, e, c.stream); });
    // l = h W2^T + b2 -> dl = (sigmoid(l) - x) / B
    e = GemmEpiArgs{};
    e.bias = P + md->off_b2; e.aux = X; e.ld_aux = Np; e.n_live = N; e.m_live = B; e.inv_batch = 1.0f / (float)B;
    e.partials = partials;
    c.launch("made_gemm", [&] { return launch_gemm(EPI_LOGITS, Bp, Np, Hp, Hh, Hp, 1, W2, 1, Hp, dL, Np, e, c.stream); });
    // weight gradients: small outputs with the batch as K, split over workgroups
    const int splits = made_splits(Bp);
    float* splitk = reinterpret_cast<float*>(c.buf("made_splitk", (size_t)splits * Np * Hp * sizeof(float)));
    // dW2 = dl^T h (masked)
    e = GemmEpiArgs{};
    e.m_live = N; e.n_live = H;
    c.launch("made_gemm", [&] {
        return launch_gemm_splitk(EPI_MASK_W2, Np, Hp, Bp, dL, 1, Np, Hh, Hp, 1, G + md->off_W2, Hp, e, splits, splitk, c.stream);
    });
    // dh = (dl W2) * [a > 0]
    e = GemmEpiArgs{};
    e.aux = Apre; e.ld_aux = Hp;
    c.launch("made_gemm", [&] { return launch_gemm(EPI_RELU_GRAD, Bp, Hp, Np, dL, Np, 1, W2, Hp, 1, dH, Hp, e, c.stream); });
    // dW1 = dh^T X (masked)
    e = GemmEpiArgs{};
    e.m_live = H; e.n_live = N;
    c.launch("made_gemm", [&] {
        return launch_gemm_splitk(EPI_MASK_W1, Hp, Np, Bp, dH, 1, Hp, X, Np, 1, G, Np, e, splits, splitk, c.stream);
    });
    float* part = reinterpret_cast<float*>(c.buf("made_colsum", (size_t)COLSUM_SLICES * (Hp > Np ? Hp : Np) * sizeof(float)));
    c.launch("made_colsum", [&] { return launch_colsum(Bp, Hp, H, dH, G + md->off_b1, part, c.stream); });
    c.launch("made_colsum", [&] { return launch_colsum(Bp, Np, N, dL, G + md->off_b2, part, c.stream); });
}

static int made_loss_tiles(const rbmhip_made* md, int Bp) { return (md->s.Np / 64) * (Bp / 64); }

// ---------------------------------------------------------------------------
// ensembles
// ---------------------------------------------------------------------------
struct rbmhip_ens {
    int kind = 0;  // 0 ExactSummation, 1 MonteCarloLoop, 2 independent MH with a MADE proposal
    rbmhip_made* made = nullptr;  // kind 2 (not owned)
    int N = 0;
    bool has_sector = false;
    int sector = 0;
    std::vector<uint64_t> configs;  // sector list, 2 words each
    int64_t num_configs = 0;
    std::map<int, DevBuf> configs_dev;
    uint32_t num_samples = 0, num_sweeps = 0, num_therm = 0, num_chains = 0;
    uint64_t seed = 0;
    uint32_t call = 0;
    uint64_t acc = 0, rej = 0;
    // MonteCarloLoop move type: has_sector selects the exchange move in the
    // sector's total S_z (set_total_z_symmetry); fast_sweep the multi-flip move
    bool fast_sweep = false;
    uint32_t fast_tries = 1;
    // persistent chains (opt-in): each call after the first starts from the
    // configurations the previous call ended in and skips thermalisation;
    // state_* identify the call that left chain_state
    bool persistent = false;
    DevBuf chain_state;
    int state_device = -1, state_N = -1, state_move = -1, state_sector = 0;
    int64_t state_chains = -1;
    // FP32 chains recompute their angles every resync_sweeps sweeps (0: never)
    int resync_sweeps = MC_RESYNC_SWEEPS;
    // final angles of the last call's chains (drift diagnostics)
    bool record_angles = false;
    DevBuf angles;
    int angles_M = 0, angles_precision = 0;
    int64_t angles_chains = 0;
    // records of the last Monte Carlo call
    int rec_device = -1;
    int64_t rec_n = 0;
    DevBuf spins, lp, el;
    // proposals of the last learned-proposal call, and its runs of equal samples
    int64_t prop_n = 0, run_n = 0;
    DevBuf prop_spins, prop_lq, prop_lp, run_mult, sample_run;
    ~rbmhip_ens() {
        for (auto& kv : configs_dev) kv.second.release();
        spins.release(); lp.release(); el.release(); chain_state.release(); angles.release();
        prop_spins.release(); prop_lq.release(); prop_lp.release(); run_mult.release(); sample_run.release();
    }
    int64_t num_steps() const { return kind == 0 ? num_configs : (int64_t)num_samples; }
};

static const uint64_t* ens_configs(Ctx& c, rbmhip_ens* e) {
    if (e->kind != 0 || !e->has_sector) return nullptr;
    DevBuf& b = e->configs_dev[c.device];
    if (b.p == nullptr) {
        b.ensure(e->configs.size() * sizeof(uint64_t));
        RBM_HIPCHK(hipMemcpy(b.p, e->configs.data(), e->configs.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    }
    return reinterpret_cast<const uint64_t*>(b.p);
}

// ---------------------------------------------------------------------------
// sampling / evaluation driver
// ---------------------------------------------------------------------------
struct Samples {
    int64_t n = 0;
    uint64_t* spins = nullptr;
    cx<double>* lp = nullptr;
    cx<double>* el = nullptr;
    void* tanh = nullptr;
    const double* mult = nullptr;   // per-configuration multiplicity (runs of a learned-proposal chain)
};

template <typename T>
static size_t rbm_w_bytes(const rbmhip_psi* psi) {
    return (size_t)psi->N * rbm_padded_width(psi->M) * sizeof(cx<T>);
}

static void check_op_fits(const rbmhip_psi* psi, const rbmhip_op* op) {
    if (op) RBM_REQUIRE(op->max_index < psi->N, "operator acts on a spin index >= N");
}

// Evaluate log psi (and E_loc / tanh) on n given configurations (configs or
// index_base + idx).
template <typename T>
static void eval_model(Ctx& c, rbmhip_psi* psi, rbmhip_op* op, int64_t n, const uint64_t* configs,
                       uint64_t index_base, const Records& rec, const char* label = "eval") {
    const OpDev od = op ? op_dev(c, op) : empty_op();
    const int max_blocks = c.num_cus * 8;
    if (psi->kind == 0) {
        const RbmDev<T> md = rbm_dev<T>(c, psi);
        const size_t lds = rbm_w_bytes<T>(psi);
        const bool wlds = lds <= c.max_lds;
        c.launch(label, [&] { return launch_eval<T>(md, od, op != nullptr, n, configs, index_base, rec, wlds, lds, max_blocks, c.stream); });
    } else if (psi->precision == RBMHIP_BF16) {
        // log psi and E_loc on the BF16 MFMA layers; O_k rows by the complex64
        // kernels on the same bfloat16 weights
        const DeepMfmaDev mm = deep_mfma_dev(c, psi);
        if (rec.spins || rec.log_psi || rec.e_loc) {
            Records r{rec.spins, rec.log_psi, rec.e_loc, nullptr};
            c.launch(label[0] == 'e' ? "deep_mfma_eval" : label, [&] {
                return launch_deep_mfma_eval(mm, od, op != nullptr, n, configs, index_base, r, c.num_cus, c.stream);
            });
        }
        if (rec.tanh) {
            const DeepDev<T> md = deep_dev<T>(c, psi);
            Records r{nullptr, nullptr, nullptr, rec.tanh};
            const OpDev none = empty_op();
            c.launch("deep_ok", [&] { return launch_deep_eval<T>(md, none, 0, n, configs, index_base, r, max_blocks, c.stream); });
        }
    } else {
        const DeepDev<T> md = deep_dev<T>(c, psi);
        c.launch(label[0] == 'e' ? "deep_eval" : label, [&] { return launch_deep_eval<T>(md, od, op != nullptr, n, configs, index_base, rec, max_blocks, c.stream); });
    }
}

// Overlap of sampling and evaluation for low-occupancy single-flip chains.
// With at most one wavefront per SIMD (16 lanes per chain) the sampler's
// per-proposal dependency chain leaves issue slots idle.  The chains then run
// in up to MC_SPLIT_CHUNKS launches, and each launch's samples but the last's
// are evaluated on the side stream while the next launch samples.  Each split
// is placed where one launch would recompute the FP32 angles from the
// configuration anyway (a multiple of the re-sync interval and of the draw
// block), so a launch that starts from the previous one's configurations with
// fresh angles and the same Philox counters (McParams::k0) reproduces one
// launch's chains bit for bit.  Returns the sample counts per chain at which
// launches end (the last is steps_per_chain), or an empty list for one
// launch.  RBMHIP_MC_SPLIT=0 in the environment turns it off.
constexpr int MC_SPLIT_CHUNKS = 8;
static std::vector<int> mc_split_points(const Ctx& c, const McParams& mp, int N, bool fp32) {
    std::vector<int> ends;
    const char* env = std::getenv("RBMHIP_MC_SPLIT");
    if (env && env[0] == '0') return ends;
    if (!fp32 || mp.move != MC_SINGLE_FLIP || mp.resync <= 0 || mp.sweeps <= 0) return ends;
    if ((int64_t)mp.chains > 16LL * c.num_cus) return ends;
    const int S = mp.steps_per_chain;
    if (S < 16) return ends;
    uint64_t L = (uint64_t)mp.resync, g = (uint64_t)MC_DRAW_BLOCK, r = L;
    while (r) { const uint64_t t = g % r; g = r; r = t; }
    L = L / g * MC_DRAW_BLOCK;
    const uint64_t base = (uint64_t)mp.therm * N, per = (uint64_t)mp.sweeps * N;
    for (int q = 1; q < MC_SPLIT_CHUNKS; q++) {
        const int target = (int)((int64_t)S * q / MC_SPLIT_CHUNKS);
        int best = 0;
        for (int s0 = (ends.empty() ? 0 : ends.back()) + 1; s0 < S; s0++)
            if ((base + (uint64_t)s0 * per) % L == 0 && (best == 0 || std::abs(s0 - target) < std::abs(best - target)))
                best = s0;
        if (best > 0 && std::abs(best - target) <= S / (2 * MC_SPLIT_CHUNKS)) ends.push_back(best);
    }
    if (ends.empty()) return ends;
    ends.push_back(S);
    return ends;
}

// Records of samples [off, ...) of a batch: per-sample pointers advanced.
template <typename T>
static Records records_from(const Records& r, int64_t off, int M) {
    Records o = r;
    if (o.spins) o.spins += 2 * off;
    if (o.log_psi) o.log_psi += off;
    if (o.e_loc) o.e_loc += off;
    if (o.tanh) o.tanh = reinterpret_cast<cx<T>*>(o.tanh) + off * (int64_t)M;
    return o;
}

// Draw the samples of an ensemble for psi and measure them.  For
// ExactSummation the samples are the basis (or sector) states.
template <typename T>
static Samples sample(Ctx& c, rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* e, bool want_spins, bool want_tanh) {
    RBM_REQUIRE(e->N == psi->N || e->kind != 0, "ensemble and wavefunction have different N");
    check_op_fits(psi, op);
    Samples s;
    s.n = e->num_steps();
    RBM_REQUIRE(s.n > 0, "the ensemble has no samples");
    // RBM: tanh(theta) records (M per sample) for the O_k contraction.  PsiDeep
    // O_k rows (num_params - 2N per sample) are rebuilt in bounded chunks from
    // the recorded configurations by contract_params instead.
    if (want_tanh && psi->kind == 0) s.tanh = c.buf("tanh", (size_t)s.n * (size_t)psi->M * sizeof(cx<T>));
    if (e->kind != 0) {
        RBM_REQUIRE(e->num_chains > 0 && e->num_samples % e->num_chains == 0,
                    "num_samples must be a positive multiple of num_markov_chains (MonteCarloLoop.cu:78)");
        e->spins.ensure((size_t)s.n * 2 * sizeof(uint64_t));
        e->lp.ensure((size_t)s.n * sizeof(cx<double>));
        e->el.ensure((size_t)s.n * sizeof(cx<double>));
        e->rec_device = c.device;
        e->rec_n = s.n;
        s.spins = reinterpret_cast<uint64_t*>(e->spins.p);
        s.lp = reinterpret_cast<cx<double>*>(e->lp.p);
        s.el = reinterpret_cast<cx<double>*>(e->el.p);
        McParams mp;
        mp.chains = (int)e->num_chains;
        mp.steps_per_chain = (int)(e->num_samples / e->num_chains);
        mp.sweeps = (int)e->num_sweeps;
        mp.therm = (int)e->num_therm;
        mp.seed_lo = (uint32_t)e->seed;
        mp.seed_hi = (uint32_t)(e->seed >> 32);
        mp.call = e->call++;
        if (e->kind == 1 && e->has_sector) {
            RBM_REQUIRE(!e->fast_sweep, "set_fast_sweep and set_total_z_symmetry cannot be combined "
                                        "(the reference's fast sweep has no exchange partner, MonteCarloLoop.hpp:160-201)");
            RBM_REQUIRE((e->sector + psi->N) % 2 == 0 && std::abs(e->sector) < psi->N,
                        "the total S_z sector must have the parity of N and |sector| < N");
            mp.move = MC_EXCHANGE;
            mp.k_up = (e->sector + psi->N) / 2;
        } else if (e->kind == 1 && e->fast_sweep) {
            mp.move = MC_FAST_SWEEP;
            mp.tries = (int)std::max<uint32_t>(1, e->fast_tries);
        }
        if (e->kind == 1 && (e->persistent || e->record_angles) && psi->kind != 0)
            throw Error(RBMHIP_ERR_UNSUPPORTED, "persistent chains and angle records are implemented for RBM wavefunctions");
        uint32_t* counts = reinterpret_cast<uint32_t*>(c.buf("counts", (size_t)mp.chains * 2 * sizeof(uint32_t)));
        uint32_t* counts_extra = nullptr;   // later launches' counts when the chains run in several
        size_t counts_extra_n = 0;
        const OpDev od = op ? op_dev(c, op) : empty_op();
        if (e->kind == 2) {
            // independent proposals do not depend on the chain: draw all of them,
            // evaluate log psi in one batch, then run the accept scan per chain
            RBM_REQUIRE(e->made && e->made->s.N == psi->N, "the proposal network and the wavefunction have different N");
            const int64_t P = (int64_t)mp.chains * (mp.therm + mp.steps_per_chain + 1);
            uint64_t* prop = reinterpret_cast<uint64_t*>(e->prop_spins.ensure((size_t)P * 2 * sizeof(uint64_t)));
            double* plq = reinterpret_cast<double*>(e->prop_lq.ensure((size_t)P * sizeof(double)));
            cx<double>* plp = reinterpret_cast<cx<double>*>(e->prop_lp.ensure((size_t)P * sizeof(cx<double>)));
            e->prop_n = P;
            const MadeInf mi = made_inf(c, e->made);
            c.launch("made_sample", [&] { return launch_made_sample(mi, P, mp.seed_lo, mp.seed_hi, mp.call, prop, plq, c.stream); });
            bool done = false;
            if constexpr (std::is_same<T, float>::value) {
                if (psi->kind == 0 && !psi->has_a) {
                    // log psi of independent configurations: one GEMM over all proposals
                    const RbmDev<float> md = rbm_dev<float>(c, psi);
                    c.launch("proposal_eval", [&] { return launch_theta_logpsi(md, P, prop, plp, c.stream); });
                    done = true;
                }
            }
            if (!done) {
                Records prec{nullptr, plp, nullptr, nullptr};
                eval_model<T>(c, psi, nullptr, P, prop, 0, prec, "proposal_eval");
            }
            // accept scan: a rejection repeats the held configuration, so each
            // chain's samples are runs of equal configurations; every run is
            // measured once and weighted by its length
            uint32_t* runs = reinterpret_cast<uint32_t*>(c.buf("imh_runs", (size_t)mp.chains * sizeof(uint32_t)));
            c.launch("imh", [&] {
                return launch_imh(mp.chains, mp.therm, mp.steps_per_chain, prop, plp, plq, mp.seed_lo, mp.seed_hi, mp.call, 0,
                                  runs, nullptr, nullptr, nullptr, nullptr, counts, c.stream);
            });
            std::vector<uint32_t> hr((size_t)mp.chains);
            c.d2h(hr.data(), runs, hr.size() * sizeof(uint32_t));
            c.sync();
            std::vector<int64_t> off((size_t)mp.chains);
            int64_t U = 0;
            for (int i = 0; i < mp.chains; i++) { off[(size_t)i] = U; U += hr[(size_t)i]; }
            int64_t* doff = reinterpret_cast<int64_t*>(c.buf("imh_off", off.size() * sizeof(int64_t)));
            c.h2d(doff, off.data(), off.size() * sizeof(int64_t));
            double* mult = reinterpret_cast<double*>(e->run_mult.ensure((size_t)U * sizeof(double)));
            int64_t* srun = reinterpret_cast<int64_t*>(e->sample_run.ensure((size_t)s.n * sizeof(int64_t)));
            c.launch("imh", [&] {
                return launch_imh(mp.chains, mp.therm, mp.steps_per_chain, prop, plp, plq, mp.seed_lo, mp.seed_hi, mp.call, 1,
                                  nullptr, doff, s.spins, mult, srun, nullptr, c.stream);
            });
            e->run_n = U;
            e->rec_n = U;
            s.n = U;
            s.mult = mult;
            Records rec{nullptr, s.lp, op ? s.el : nullptr, s.tanh};
            eval_model<T>(c, psi, op, s.n, s.spins, 0, rec);
        } else if (psi->kind == 0) {
            // sampling records configurations; measurement runs at full
            // occupancy in the evaluation kernel on the recorded samples
            const RbmDev<T> md = rbm_dev<T>(c, psi);
            const size_t lds = rbm_w_bytes<T>(psi);
            const bool wlds = lds + 4096 <= c.max_lds;
            mp.resync = e->resync_sweeps * psi->N;
            if (e->persistent) {
                const bool same = e->state_device == c.device && e->state_chains == mp.chains &&
                                  e->state_N == psi->N && e->state_move == mp.move &&
                                  (mp.move != MC_EXCHANGE || e->state_sector == e->sector);
                uint64_t* st = reinterpret_cast<uint64_t*>(e->chain_state.ensure((size_t)mp.chains * 2 * sizeof(uint64_t)));
                if (same) {
                    mp.start = st;
                    mp.therm = 0;
                }
                mp.final_out = st;
            }
            if (e->record_angles) {
                mp.angles_out = e->angles.ensure((size_t)mp.chains * psi->M * sizeof(cx<T>));
                e->angles_M = psi->M;
                e->angles_chains = mp.chains;
                e->angles_precision = sizeof(T) == 4 ? RBMHIP_FP32 : RBMHIP_FP64;
            }
            Records rec{nullptr, s.lp, op ? s.el : nullptr, s.tanh};
            const std::vector<int> ends = mc_split_points(c, mp, psi->N, sizeof(T) == 4);
            if (!ends.empty()) {
                c.ensure_side();
                const size_t K = ends.size();
                uint64_t* st = mp.final_out ? mp.final_out
                                            : reinterpret_cast<uint64_t*>(c.buf("mc_split_state", (size_t)mp.chains * 2 * sizeof(uint64_t)));
                counts_extra = reinterpret_cast<uint32_t*>(c.buf("mc_split_counts", (K - 1) * mp.chains * 2 * sizeof(uint32_t)));
                counts_extra_n = (K - 1) * mp.chains * 2;
                int begin = 0;
                for (size_t q = 0; q < K; q++) {
                    const int64_t off = (int64_t)begin * mp.chains, cnt = (int64_t)(ends[q] - begin) * mp.chains;
                    McParams a = mp;
                    a.steps_per_chain = ends[q] - begin;
                    if (q > 0) {
                        a.therm = 0;
                        a.start = st;
                        a.k0 = (uint64_t)mp.therm * psi->N + (uint64_t)begin * mp.sweeps * psi->N;
                    }
                    if (q + 1 < K) {
                        a.final_out = st;
                        a.angles_out = nullptr;
                    }
                    uint32_t* cq = q == 0 ? counts : counts_extra + (q - 1) * mp.chains * 2;
                    c.launch("mc", [&] { return launch_mc<T>(md, a, s.spins + 2 * off, cq, wlds, lds, c.stream); });
                    const Records rq = records_from<T>(rec, off, psi->M);
                    if (q + 1 < K) {
                        // evaluate beside the next launch
                        RBM_HIPCHK(hipEventRecord(c.ev_main, c.stream));
                        RBM_HIPCHK(hipStreamWaitEvent(c.side, c.ev_main, 0));
                        std::swap(c.stream, c.side);
                        try {
                            eval_model<T>(c, psi, op, cnt, s.spins + 2 * off, 0, rq);
                        } catch (...) {
                            std::swap(c.stream, c.side);
                            throw;
                        }
                        std::swap(c.stream, c.side);
                    } else {
                        eval_model<T>(c, psi, op, cnt, s.spins + 2 * off, 0, rq);
                    }
                    begin = ends[q];
                }
                RBM_HIPCHK(hipEventRecord(c.ev_side, c.side));
                RBM_HIPCHK(hipStreamWaitEvent(c.stream, c.ev_side, 0));
            } else {
                if (mp.move == MC_SINGLE_FLIP)
                    c.launch("mc", [&] { return launch_mc<T>(md, mp, s.spins, counts, wlds, lds, c.stream); });
                else
                    c.launch("mc", [&] { return launch_mc_moves<T>(md, mp, s.spins, counts, c.stream); });
                eval_model<T>(c, psi, op, s.n, s.spins, 0, rec);
            }
            if (e->persistent) {
                e->state_device = c.device;
                e->state_chains = mp.chains;
                e->state_N = psi->N;
                e->state_move = mp.move;
                e->state_sector = e->sector;
            }
        } else if (psi->precision == RBMHIP_BF16) {
            const DeepMfmaDev mm = deep_mfma_dev(c, psi);
            c.launch("deep_mfma_mc", [&] { return launch_deep_mfma_mc(mm, mp, s.spins, counts, c.num_cus, c.stream); });
            Records rec{nullptr, s.lp, op ? s.el : nullptr, s.tanh};
            eval_model<T>(c, psi, op, s.n, s.spins, 0, rec);
        } else {
            Records rec{s.spins, s.lp, op ? s.el : nullptr, s.tanh};
            const DeepDev<T> md = deep_dev<T>(c, psi);
            c.launch("deep_mc", [&] { return launch_deep_mc<T>(md, od, op != nullptr, mp, rec, counts, c.stream); });
        }
        std::vector<uint32_t> h((size_t)mp.chains * 2), h2;
        c.d2h(h.data(), counts, h.size() * sizeof(uint32_t));
        if (counts_extra) {
            h2.resize(counts_extra_n);
            c.d2h(h2.data(), counts_extra, h2.size() * sizeof(uint32_t));
        }
        c.sync();
        e->acc = e->rej = 0;
        for (int i = 0; i < mp.chains; i++) { e->acc += h[2 * i]; e->rej += h[2 * i + 1]; }
        for (size_t i = 0; i + 1 < h2.size(); i += 2) { e->acc += h2[i]; e->rej += h2[i + 1]; }
        if (!op) s.el = nullptr;
    } else {
        s.lp = reinterpret_cast<cx<double>*>(c.buf("lp", (size_t)s.n * sizeof(cx<double>)));
        if (op) s.el = reinterpret_cast<cx<double>*>(c.buf("el", (size_t)s.n * sizeof(cx<double>)));
        if (want_spins) s.spins = reinterpret_cast<uint64_t*>(c.buf("spins", (size_t)s.n * 2 * sizeof(uint64_t)));
        Records rec{s.spins, s.lp, s.el, s.tanh};
        eval_model<T>(c, psi, op, s.n, ens_configs(c, e), 0, rec);
    }
    return s;
}

struct Scalars {
    double v[SC_STRIDE];
};

static Scalars reduce_scalars(Ctx& c, int64_t n, const double* sc) {
    int nb = (int)std::min<int64_t>(1024, (n + 255) / 256);
    if (nb < 1) nb = 1;
    double* part = reinterpret_cast<double*>(c.buf("scpart", (size_t)nb * SC_STRIDE * sizeof(double)));
    c.launch("scalar_reduce", [&] { return launch_scalar_partials(n, sc, nb, part, c.stream); });
    std::vector<double> h((size_t)nb * SC_STRIDE);
    c.d2h(h.data(), part, h.size() * sizeof(double));
    c.sync();
    Scalars r;
    for (int q = 0; q < SC_STRIDE; q++) {
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += h[(size_t)b * SC_STRIDE + q];
        r.v[q] = s;
    }
    return r;
}

// R[r][col] = sum_s coef[s] row_r(s) conj(col(s)) for the RBM rows
// {1, s_0..s_{N-1}} and columns {T_0..T_{M-1}, 1}: FP32 arithmetic on the
// bfloat16 matrix cores (exact three-part split) for complex64 records, chunk
// partials added in FP64; FP64 VALU on complex128 records.
template <typename T>
static std::vector<cd> ok_contract(Ctx& c, int N, int M, int64_t n, const void* tanh, const uint64_t* spins,
                                   const cx<double>* coef, bool fp64_sums = false) {
    const int64_t plen = (int64_t)(N + 1) * (M + 1);
    cx<double>* out = reinterpret_cast<cx<double>*>(c.buf("okout", (size_t)plen * sizeof(cx<double>)));
    if (sizeof(T) == sizeof(float) && !fp64_sums) {
        int rows_p = 0, cols_p = 0;
        ok_mfma_shape(N, M, &rows_p, &cols_p);
        const int64_t per_chunk = (int64_t)rows_p * cols_p;
        int64_t nch = std::min<int64_t>(2048, std::max<int64_t>(1, (256ll << 20) / (per_chunk * 4)));
        nch = std::min<int64_t>(nch, std::max<int64_t>(1, (n + 63) / 64));
        int64_t chunk = (n + nch - 1) / nch;
        chunk = (chunk + 31) / 32 * 32;   // whole 32-sample K-steps
        nch = (n + chunk - 1) / chunk;
        float* part = reinterpret_cast<float*>(c.buf("mfpart", (size_t)(nch * per_chunk) * sizeof(float)));
        double* psum = reinterpret_cast<double*>(c.buf("mfsum", (size_t)per_chunk * MFMA_SUM_SPLIT * sizeof(double)));
        uint32_t* planes = ok_mfma_needs_planes(N)
                               ? reinterpret_cast<uint32_t*>(c.buf("okplanes", (size_t)((n + 31) / 32) * N * sizeof(uint32_t)))
                               : nullptr;
        c.launch("ok_reduce", [&] {
            return launch_ok_mfma(n, N, M, tanh, spins, coef, chunk, (int)nch, rows_p, cols_p, part, planes, c.stream);
        });
        c.launch("partials_sum", [&] { return launch_mfma_finish((int)nch, N, M, rows_p, cols_p, part, psum, out, c.stream); });
    } else {
        int64_t nb = std::min<int64_t>(1024, std::max<int64_t>(1, (256ll << 20) / (plen * 16)));
        nb = std::min<int64_t>(nb, std::max<int64_t>(1, (n + 63) / 64));
        const int64_t chunk = (n + nb - 1) / nb;
        cx<double>* part = reinterpret_cast<cx<double>*>(c.buf("okpart", (size_t)(nb * plen) * sizeof(cx<double>)));
        c.launch("ok_reduce", [&] { return launch_ok_reduce<T>(n, N, M, tanh, spins, coef, 1, chunk, (int)nb, part, c.stream); });
        c.launch("partials_sum", [&] { return launch_partials_sum((int)nb, plen, part, out, c.stream); });
    }
    std::vector<cd> h((size_t)plen);
    c.d2h(h.data(), out, h.size() * sizeof(cx<double>));
    c.sync();
    return h;
}

// BF16 PsiDeep: sum_s coef(s) conj(O_k(s)) in parameter order from gradient
// records (x_l, delta_l, log psi per configuration, MFMA backward pass) and one
// FP32 MFMA GEMM per layer over the samples, G_l = sum_s x_{l-1}(s) (x)
// conj(coef_s) delta_l(s) on real rows, split over K and accumulated in FP64:
// the O_k rows (num_params - 2N per sample) are never formed.
static bool deep_mfma_grad_fits(Ctx& c, rbmhip_psi* psi) {
    if (!deep_mfma_fits(psi, nullptr)) return false;
    const DeepMfmaDev md = deep_mfma_dev(c, psi);
    return deep_mfma_lds_bytes(md, true) <= c.max_lds;
}

static std::vector<cd> deep_mfma_contract(Ctx& c, rbmhip_psi* psi, int64_t n, const uint64_t* spins, const cx<double>* coef) {
    RBM_REQUIRE(spins != nullptr, "PsiDeep O_k needs the recorded configurations");
    const DeepMfmaDev md = deep_mfma_dev(c, psi);
    const int L = md.L, N = psi->N;
    // 256 K-splits: the per-layer products have 2 to 6 output tiles of 128 x 128
    const int SPLITS = 256, KQ = 16 * SPLITS;
    auto pad128 = [](int x) { return (x + 127) / 128 * 128; };   // 128 x 128 MFMA tiles
    std::vector<int> in(L + 1), out(L + 1), ldx(L + 1), ldy(L + 1);
    int mx = 0, my = 0;
    for (int l = 0; l <= L; l++) {
        in[l] = l == 0 ? N : psi->sizes[l - 1];
        out[l] = l < L ? psi->sizes[l] : 1;
        ldx[l] = pad128(2 * in[l] + (l < L ? 1 : 0));
        ldy[l] = pad128(2 * out[l]);
        mx = std::max(mx, ldx[l]);
        my = std::max(my, ldy[l]);
    }
    const int64_t chunk = std::min<int64_t>(n, 1 << 18);
    const int64_t rows_max = (chunk + KQ - 1) / KQ * KQ;
    cx<float>* rec = reinterpret_cast<cx<float>*>(c.buf("dg_records", (size_t)chunk * md.rec_stride * sizeof(cx<float>)));
    float* X = reinterpret_cast<float*>(c.buf("dg_x", (size_t)rows_max * mx * sizeof(float)));
    float* Y = reinterpret_cast<float*>(c.buf("dg_y", (size_t)rows_max * my * sizeof(float)));
    float* part = reinterpret_cast<float*>(c.buf("dg_part", (size_t)SPLITS * mx * my * sizeof(float)));
    std::vector<size_t> goff(L + 2, 0);
    for (int l = 0; l <= L; l++) goff[l + 1] = goff[l] + (size_t)ldx[l] * ldy[l];
    double* Gd = reinterpret_cast<double*>(c.buf("dg_g64", goff[L + 1] * sizeof(double)));
    c.launch("deep_grad", [&] { return launch_zero((int64_t)(2 * goff[L + 1]), reinterpret_cast<float*>(Gd), c.stream); });
    for (int64_t s0 = 0; s0 < n; s0 += chunk) {
        const int64_t cnt = std::min<int64_t>(chunk, n - s0), rows = (cnt + KQ - 1) / KQ * KQ;
        c.launch("deep_grad", [&] { return launch_deep_mfma_grad(md, cnt, spins + 2 * s0, rec, c.num_cus, c.stream); });
        for (int l = 0; l <= L; l++) {
            c.launch("deep_grad_rows", [&] {
                return launch_deep_grad_rows(md, l, in[l], out[l], 0, cnt, rows, spins + 2 * s0, rec, coef + s0,
                                             psi->ref_compat_final, X, ldx[l], Y, ldy[l], c.stream);
            });
            c.launch("deep_grad_gemm", [&] {
                return launch_gemm_tn_splitk(ldx[l], ldy[l], rows, SPLITS, X, Y, part, c.stream);
            });
            c.launch("deep_grad_gemm", [&] {
                return launch_accum_splits_f64((int64_t)ldx[l] * ldy[l], SPLITS, part, Gd + goff[l], c.stream);
            });
        }
    }
    std::vector<double> h(goff[L + 1]);
    c.d2h(h.data(), Gd, h.size() * sizeof(double));
    c.sync();
    std::vector<cd> one((size_t)psi->num_params(), cd(0, 0));
    int64_t p = 2 * (int64_t)N, woff = 0;
    for (int l = 0; l < L; l++) {
        const double* g = h.data() + goff[l];
        const int I = in[l], J = out[l], ly = ldy[l];
        for (int j = 0; j < J; j++) one[(size_t)(p + j)] = std::conj(cd(g[(size_t)(2 * I) * ly + j], g[(size_t)(2 * I) * ly + J + j]));
        p += J;
        const int C = psi->conn_counts[l];
        for (int cc = 0; cc < C; cc++)
            for (int j = 0; j < J; j++) {
                const int i = (int)psi->conns[(size_t)(woff + (int64_t)cc * J + j)];
                const double rr = g[(size_t)i * ly + j], ii = g[(size_t)(I + i) * ly + J + j];
                const double ri = g[(size_t)i * ly + J + j], ir = g[(size_t)(I + i) * ly + j];
                one[(size_t)(p + (int64_t)cc * J + j)] = std::conj(cd(rr - ii, ri + ir));
            }
        p += (int64_t)C * J;
        woff += (int64_t)C * J;
    }
    {
        const double* g = h.data() + goff[L];
        const int I = in[L], ly = ldy[L];
        for (int j = 0; j < I; j++) {
            const double rr = g[(size_t)j * ly], ri = g[(size_t)j * ly + 1];
            const double ir = g[(size_t)(I + j) * ly], ii = g[(size_t)(I + j) * ly + 1];
            one[(size_t)(p + j)] = std::conj(cd(rr - ii, ri + ir));
        }
    }
    return one;
}

// Map the contraction R of one coefficient vector to parameter order
// [alpha (N), beta (N), b (M), W (N x M), a (N)]:  entries are
// sum_s coef_s conj(O_k(s)) (Psi.hpp:165-202).
static void rbm_params_from_R(const rbmhip_psi* psi, const cd* R, std::vector<cd>& out) {
    const int N = psi->N, M = psi->M, C = M + 1;
    out.assign((size_t)psi->num_params(), cd(0, 0));
    for (int j = 0; j < M; j++) out[2 * N + j] = R[j];
    for (int i = 0; i < N; i++)
        for (int j = 0; j < M; j++) out[2 * N + M + (size_t)i * M + j] = R[(size_t)(i + 1) * C + j];
    if (psi->has_a)
        for (int i = 0; i < N; i++) out[2 * N + M + (size_t)N * M + i] = R[(size_t)(i + 1) * C + M];
}

// sum_s coef(s) conj(O_k(s)) in parameter order (entries below 2N are zero).
template <typename T>
static std::vector<cd> contract_params(Ctx& c, rbmhip_psi* psi, int64_t n, const void* tanh, const uint64_t* spins,
                                       const cx<double>* coef, bool fp64_sums = false) {
    std::vector<cd> one;
    if (psi->kind == 0) {
        std::vector<cd> R = ok_contract<T>(c, psi->N, psi->M, n, tanh, spins, coef, fp64_sums);
        rbm_params_from_R(psi, R.data(), one);
        return one;
    }
    if (psi->precision == RBMHIP_BF16 && deep_mfma_grad_fits(c, psi)) return deep_mfma_contract(c, psi, n, spins, coef);
    // PsiDeep: O_k rows of a chunk of the recorded configurations are
    // materialised by the deep kernels and reduced column-wise; chunks keep the
    // row buffer at most 4 GiB and their FP64 sums are added on the host.
    RBM_REQUIRE(spins != nullptr, "PsiDeep O_k needs the recorded configurations");
    (void)tanh;
    const int64_t K = psi->num_params() - 2 * (int64_t)psi->N;
    const int64_t rows_max = std::max<int64_t>(64, std::min<int64_t>(n, (4ll << 30) / (K * (int64_t)sizeof(cx<T>))));
    void* rows = c.buf("deep_ok_rows", (size_t)rows_max * K * sizeof(cx<T>));
    cx<double>* out = reinterpret_cast<cx<double>*>(c.buf("okout", (size_t)K * sizeof(cx<double>)));
    std::vector<cd> acc((size_t)K, cd(0, 0)), h((size_t)K);
    for (int64_t s0 = 0; s0 < n; s0 += rows_max) {
        const int64_t cur = std::min<int64_t>(rows_max, n - s0);
        Records rec{nullptr, nullptr, nullptr, rows};
        eval_model<T>(c, psi, nullptr, cur, spins + 2 * s0, 0, rec, "deep_ok");
        int64_t nb = std::min<int64_t>(1024, std::max<int64_t>(1, (256ll << 20) / (K * 16)));
        nb = std::min<int64_t>(nb, std::max<int64_t>(1, (cur + 63) / 64));
        const int64_t chunk = (cur + nb - 1) / nb;
        cx<double>* part = reinterpret_cast<cx<double>*>(c.buf("okpart", (size_t)(nb * K) * sizeof(cx<double>)));
        c.launch("ok_reduce", [&] { return launch_rows_reduce<T>(cur, K, rows, coef + s0, 1, chunk, (int)nb, part, c.stream); });
        c.launch("partials_sum", [&] { return launch_partials_sum((int)nb, K, part, out, c.stream); });
        c.d2h(h.data(), out, h.size() * sizeof(cx<double>));
        c.sync();
        for (int64_t k = 0; k < K; k++) acc[(size_t)k] += h[(size_t)k];
    }
    one.assign((size_t)psi->num_params(), cd(0, 0));
    for (int64_t k = 0; k < K; k++) one[(size_t)(2 * psi->N + k)] = acc[(size_t)k];
    return one;
}

template <class F>
static void with_precision(int precision, F&& f) {
    if (precision == RBMHIP_FP64) f(double(0)); else f(float(0));
}

static void check_psi(const rbmhip_psi* p) { RBM_REQUIRE(p != nullptr, "null wavefunction handle"); }
static void check_ens(const rbmhip_ens* e) { RBM_REQUIRE(e != nullptr, "null ensemble handle"); }

template <typename T>
static Scalars energy_scalars(Ctx& c, rbmhip_psi* psi, const Samples& s, rbmhip_ens* ens, cx<double>* coef) {
    double* sc = reinterpret_cast<double*>(c.buf("sc", (size_t)s.n * SC_STRIDE * sizeof(double)));
    CoefArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = s.n; a.mode = 0; a.exact = ens->kind == 0; a.num_configs = (double)ens->num_configs; a.mult = s.mult;
    a.log_prefactor = std::log(psi->prefactor); a.lp = s.lp; a.el = s.el; a.coef = coef; a.sc = sc;
    c.launch("coef", [&] { return launch_coef(a, c.stream); });
    return reduce_scalars(c, s.n, sc);
}

// Allocation-free local energies: rbmhip_local_energy with every device
// buffer it needs (the model's padded W, b and a, the operator's canonical
// tables, and the spins / log psi / E_loc records) placed in this device's
// scratch array instead of library allocations.  Fixed layout: tables in
// [0, NOALLOC_TABLE_BYTES), then three regions of NOALLOC_MAX_CONFIGS x 16
// bytes.  Nothing here needs clearing: every byte a kernel reads is written
// by this call first (the padded W columns are uploaded as zeros).
namespace {
struct ScratchTables {
    unsigned char* base;
    size_t used = 0;
    template <class V>
    void* put(Ctx& c, const V& v) {
        const size_t bytes = v.size() * sizeof(typename V::value_type);
        const size_t at = (used + 255) & ~(size_t)255;
        RBM_REQUIRE(at + bytes <= NOALLOC_TABLE_BYTES,
                    "the model and operator need more than the " + std::to_string(NOALLOC_TABLE_BYTES >> 20) +
                        " MiB of table space of the allocation-free entry point");
        used = at + bytes;
        if (bytes) c.h2d(base + at, v.data(), bytes);
        return base + at;
    }
};
}  // namespace

template <typename T>
static void local_energy_noalloc_impl(Ctx& c, rbmhip_psi* psi, rbmhip_op* op, const uint64_t* spins, int64_t n,
                                      double* out) {
    unsigned char* base = reinterpret_cast<unsigned char*>(c.scratch);
    ScratchTables tab{base};
    // model: W padded with zero columns to ld hidden units, b to ld, a (N)
    const int N = psi->N, M = psi->M, ld = rbm_padded_width(M);
    std::vector<cx<T>> Wp((size_t)N * ld, mk<T>(0, 0)), bp((size_t)ld, mk<T>(0, 0)), ap;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < M; j++) {
            const cd w = psi->W[(size_t)i * M + j];
            Wp[(size_t)i * ld + j] = mk<T>((T)w.real(), (T)w.imag());
        }
    for (int j = 0; j < M; j++) bp[j] = mk<T>((T)psi->b[j].real(), (T)psi->b[j].imag());
    if (psi->has_a)
        for (const cd& v : psi->a) ap.push_back(mk<T>((T)v.real(), (T)v.imag()));
    RbmDev<T> md;
    md.N = N;
    md.M = M;
    md.ld = ld;
    md.W = reinterpret_cast<const cx<T>*>(tab.put(c, Wp));
    md.b = reinterpret_cast<const cx<T>*>(tab.put(c, bp));
    md.a = psi->has_a ? reinterpret_cast<const cx<T>*>(tab.put(c, ap)) : nullptr;
    // operator tables in the canonical form of canonicalise()
    OpDev od;
    od.n_diag = (int)op->diag_coef.size();
    od.diag_coef = reinterpret_cast<const cx<double>*>(tab.put(c, op->diag_coef));
    od.diag_zy = reinterpret_cast<const uint64_t*>(tab.put(c, op->diag_zy));
    od.n_terms = (int)op->flip_begin.size() - 1;
    od.flip_begin = reinterpret_cast<const int*>(tab.put(c, op->flip_begin));
    od.flips = reinterpret_cast<const int*>(tab.put(c, op->flips));
    od.ent_begin = reinterpret_cast<const int*>(tab.put(c, op->ent_begin));
    od.ent_coef = reinterpret_cast<const cx<double>*>(tab.put(c, op->ent_coef));
    od.ent_zy = reinterpret_cast<const uint64_t*>(tab.put(c, op->ent_zy));
    // records
    const size_t region = (size_t)NOALLOC_MAX_CONFIGS * 16;
    uint64_t* ds = reinterpret_cast<uint64_t*>(base + NOALLOC_TABLE_BYTES);
    cx<double>* lp = reinterpret_cast<cx<double>*>(base + NOALLOC_TABLE_BYTES + region);
    cx<double>* el = reinterpret_cast<cx<double>*>(base + NOALLOC_TABLE_BYTES + 2 * region);
    c.h2d(ds, spins, (size_t)n * 2 * sizeof(uint64_t));
    const size_t lds = rbm_w_bytes<T>(psi);
    Records rec{nullptr, lp, el, nullptr};
    RBM_HIPCHK(launch_eval<T>(md, od, 1, n, ds, 0, rec, lds <= c.max_lds, lds, c.num_cus * 8, c.stream));
    c.d2h(out, el, (size_t)n * sizeof(cx<double>));
    // the host staging vectors above stay alive until the copies are done
    RBM_HIPCHK(hipStreamSynchronize(c.stream));
}

// ===========================================================================
// C ABI
// ===========================================================================
// Centred, weighted S over the parameters with an O_k (the `params` entries
// after alpha and beta), FP32 MFMA Gram blocks accumulated in FP64; returns a
// device array of P x P complex128 (eps added on the diagonal).
template <typename T>
static cx<double>* s_matrix_centred(Ctx& c, rbmhip_psi* psi, rbmhip_ens* ens, const Samples& s, double eps) {
    RBM_REQUIRE(psi->kind == 0, "the centred S-matrix is built for Psi");
    const int N = psi->N, M = psi->M;
    const int P = (int)(psi->num_params() - 2 * (int64_t)N);
    const double n = (double)ens->num_steps();
    cx<double>* w = reinterpret_cast<cx<double>*>(c.buf("s_w", (size_t)s.n * sizeof(cx<double>)));
    CoefArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = s.n; a.mode = 4; a.exact = ens->kind == 0; a.num_configs = (double)ens->num_configs; a.mult = s.mult;
    a.log_prefactor = std::log(psi->prefactor); a.lp = s.lp; a.coef = w;
    c.launch("coef", [&] { return launch_coef(a, c.stream); });
    const std::vector<cd> osum = contract_params<T>(c, psi, s.n, s.tanh, s.spins, w);
    std::vector<cx<float>> obar((size_t)P);
    for (int k = 0; k < P; k++) {
        // the contraction gives sum_s w conj(O_k)
        const cd o = std::conj(osum[(size_t)(2 * N + k)]) / n;
        obar[(size_t)k] = mk((float)o.real(), (float)o.imag());
    }
    cx<float>* dob = reinterpret_cast<cx<float>*>(c.buf("s_obar", (size_t)P * sizeof(cx<float>)));
    c.h2d(dob, obar.data(), obar.size() * sizeof(cx<float>));
    const int ld2 = (2 * P + 127) / 128 * 128;   // k_gram_lower_bf3 works on 128 x 128 tiles
    // rows of a batch: 6 bytes per value (three bfloat16 parts), at most 768 MiB
    int64_t nb = ((int64_t)(768ll << 20) / ((int64_t)ld2 * 6)) / 64 * 64;
    nb = std::max<int64_t>(64, std::min<int64_t>(nb, (s.n + 63) / 64 * 64));
    uint32_t* O = reinterpret_cast<uint32_t*>(c.buf("s_rows", (size_t)nb * ld2 * 6));
    double* Gd = reinterpret_cast<double*>(c.buf("s_gram64", (size_t)ld2 * ld2 * sizeof(double)));
    std::vector<int2> order;
    gram_tile_order(ld2, order);
    int2* dorder = reinterpret_cast<int2*>(c.buf("s_gram_order", order.size() * sizeof(int2)));
    c.h2d(dorder, order.data(), order.size() * sizeof(int2));
    c.launch("s_matrix", [&] { return launch_zero((int64_t)2 * ld2 * ld2, reinterpret_cast<float*>(Gd), c.stream); });
    for (int64_t s0 = 0; s0 < s.n; s0 += nb) {
        const int64_t cur = std::min<int64_t>(nb, s.n - s0), curp = (cur + 63) / 64 * 64;
        c.launch("s_matrix", [&] {
            return launch_o_split3<T>(s0, cur, curp, N, M, P, ld2, s.tanh, s.spins, w, dob, O, c.stream);
        });
        c.launch("s_gram", [&] {
            return launch_gram_lower(ld2, (int)curp, O, Gd, dorder, (int)order.size(), c.stream);
        });
    }
    cx<double>* S = reinterpret_cast<cx<double>*>(c.buf("s_centred", (size_t)P * P * sizeof(cx<double>)));
    c.launch("s_matrix", [&] { return launch_s_assemble(P, ld2, 1.0 / n, eps, Gd, S, c.stream); });
    return S;
}

// Cholesky-factor the device matrix A (n x n, lower triangle) and solve A x = x.
static void cholesky_solve_device(Ctx& c, int n, cx<double>* A, cx<double>* x) {
    int* info = reinterpret_cast<int*>(c.buf("chol_info", 64));
    c.launch("cholesky", [&] { return launch_zero(16, reinterpret_cast<float*>(info), c.stream); });
    c.launch("cholesky", [&] { return launch_cholesky(n, A, info, c.stream); });
    int h = 0;
    c.d2h(&h, info, sizeof(int));
    c.sync();
    if (h) throw Error(RBMHIP_ERR_NUMERIC, "matrix is not positive definite (pivot " + std::to_string(h) + ")");
    c.launch("cholesky_solve", [&] { return launch_cholesky_solve(n, A, x, c.stream); });
}

extern "C" {

const char* rbmhip_last_error(void) { return g_err.c_str(); }
int rbmhip_version(void) { return 10000; }

int rbmhip_set_device(int device) {
    return guard([&] {
        RBM_REQUIRE(device >= 0, "device index must be >= 0");
        std::lock_guard<std::mutex> lk(g_mu);
        g_device = device;
    });
}

int rbmhip_device_count(int* count) {
    return guard([&] {
        RBM_REQUIRE(count, "null output");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

// cudaProfilerStart / cudaProfilerStop (source/RBMonGPU.cu:11-17) become the
// ROCm tools' collection switches: rocprofv3 --selected-regions collects only
// between roctxProfilerResume and roctxProfilerPause, and the named range
// marks the region in traces.
int rbmhip_start_profiling(void) {
    return guard([&] {
        Ctx& c = ctx();
        c.sync();
        roctxProfilerResume(0);
        roctxRangePush("rbmhip profiled region");
    });
}
int rbmhip_stop_profiling(void) {
    return guard([&] {
        Ctx& c = ctx();
        c.sync();
        roctxRangePop();
        roctxProfilerPause(0);
    });
}

int rbmhip_activation(const double* z, int64_t n, int derivative, double* out) {
    return guard([&] {
        RBM_REQUIRE(n >= 0 && (n == 0 || (z && out)), "bad arguments");
        if (n == 0) return;
        Ctx& c = ctx();
        cx<double>* dz = reinterpret_cast<cx<double>*>(c.buf("act_in", (size_t)n * sizeof(cx<double>)));
        cx<double>* dout = reinterpret_cast<cx<double>*>(c.buf("act_out", (size_t)n * sizeof(cx<double>)));
        c.h2d(dz, z, (size_t)n * sizeof(cx<double>));
        c.launch("activation", [&] { return launch_activation(n, dz, dout, derivative, c.stream); });
        c.d2h(out, dout, (size_t)n * sizeof(cx<double>));
        c.sync();
    });
}

// ---- wavefunctions --------------------------------------------------------
static void fill_cd(std::vector<cd>& v, const double* src, size_t n) {
    v.resize(n);
    for (size_t i = 0; i < n; i++) v[i] = cd(src[2 * i], src[2 * i + 1]);
}

int rbmhip_rbm_create(int N, int M, const double* b, const double* W, const double* a, double prefactor, int precision,
                      rbmhip_psi** out) {
    return guard([&] {
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(N >= 1 && N <= 128, "N must be in [1, 128]");
        RBM_REQUIRE(M >= 1 && M <= max_hidden_units(), "M must be in [1, 256]");
        RBM_REQUIRE(b && W, "b and W are required");
        RBM_REQUIRE(precision == RBMHIP_FP32 || precision == RBMHIP_FP64, "precision must be 32 or 64");
        auto p = std::make_unique<rbmhip_psi>();
        p->kind = 0; p->N = N; p->M = M; p->precision = precision; p->prefactor = prefactor;
        fill_cd(p->b, b, (size_t)M);
        fill_cd(p->W, W, (size_t)N * M);
        if (a) { p->has_a = true; fill_cd(p->a, a, (size_t)N); }
        *out = p.release();
    });
}

int rbmhip_rbm_set_params(rbmhip_psi* psi, const double* b, const double* W, const double* a) {
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(psi->kind == 0, "not an RBM");
        if (b) fill_cd(psi->b, b, (size_t)psi->M);
        if (W) fill_cd(psi->W, W, (size_t)psi->N * psi->M);
        if (a) { RBM_REQUIRE(psi->has_a, "this RBM has no visible bias"); fill_cd(psi->a, a, (size_t)psi->N); }
        psi->version++;
    });
}

int rbmhip_deep_create(int N, int num_layers, const int32_t* sizes, const int32_t* conn_counts, const uint32_t* conns,
                       const double* weights, const double* biases, const double* final_weights, double prefactor,
                       int translational_invariance, int ref_compat_final, int precision, rbmhip_psi** out) {
    return guard([&] {
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(N >= 1 && N <= 128, "N must be in [1, 128]");
        RBM_REQUIRE(num_layers >= 1 && num_layers <= DEEP_MAX_LAYERS, "1 to 4 layers are supported");
        RBM_REQUIRE(sizes && conn_counts && conns && weights && biases && final_weights, "null layer tables");
        RBM_REQUIRE(precision == RBMHIP_FP32 || precision == RBMHIP_FP64 || precision == RBMHIP_BF16,
                    "precision must be 32, 64 or 16 (BF16 MFMA layers)");
        auto p = std::make_unique<rbmhip_psi>();
        p->kind = 1; p->N = N; p->precision = precision; p->prefactor = prefactor;
        p->translational_invariance = translational_invariance ? 1 : 0;
        p->ref_compat_final = ref_compat_final ? 1 : 0;
        int64_t nw = 0, nb = 0;
        int prev = N, units = 0;
        for (int l = 0; l < num_layers; l++) {
            RBM_REQUIRE(sizes[l] >= 1 && sizes[l] <= DEEP_MAX_WIDTH, "layer width must be in [1, 256]");
            RBM_REQUIRE(conn_counts[l] >= 1 && conn_counts[l] <= prev, "connectivity must be in [1, previous width]");
            p->sizes.push_back(sizes[l]);
            p->conn_counts.push_back(conn_counts[l]);
            for (int64_t k = 0; k < (int64_t)conn_counts[l] * sizes[l]; k++) {
                const uint32_t u = conns[nw + k];
                RBM_REQUIRE((int)u < prev, "connection index out of range of the previous layer");
            }
            nw += (int64_t)conn_counts[l] * sizes[l];
            nb += sizes[l];
            prev = sizes[l];
            units += sizes[l];
        }
        RBM_REQUIRE(units <= DEEP_MAX_UNITS, "too many units in total");
        p->conns.assign(conns, conns + nw);
        // Transpose of layer l+1's connections (the reference's rhs tables,
        // source/quantum_state/PsiDeep.cu:86-117, same visiting order): for unit
        // i of layer l, the slots c * size_{l+1} + j of layer l+1 that read it.
        int64_t woff = 0;
        for (int l = 0; l < num_layers; l++) {
            const int64_t next_off = woff + (int64_t)conn_counts[l] * sizes[l];
            p->rb_off.push_back((int)p->rhs_begin.size());
            if (l + 1 < num_layers) {
                const int size = sizes[l], nsize = sizes[l + 1], nC = conn_counts[l + 1];
                std::vector<std::vector<int>> lists((size_t)size);
                for (int j = 0; j < nsize; j++)
                    for (int c = 0; c < nC; c++)
                        lists[conns[next_off + (int64_t)c * nsize + j]].push_back(c * nsize + j);
                for (int i = 0; i < size; i++) {
                    p->rhs_begin.push_back((int)p->rhs_entries.size());
                    p->rhs_entries.insert(p->rhs_entries.end(), lists[i].begin(), lists[i].end());
                }
                p->rhs_begin.push_back((int)p->rhs_entries.size());
            }
            woff = next_off;
        }
        fill_cd(p->weights, weights, (size_t)nw);
        fill_cd(p->biases, biases, (size_t)nb);
        fill_cd(p->final_weights, final_weights, (size_t)sizes[num_layers - 1]);
        *out = p.release();
    });
}

int rbmhip_deep_set_params(rbmhip_psi* psi, const double* weights, const double* biases, const double* final_weights) {
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(psi->kind == 1, "not a PsiDeep");
        if (weights) fill_cd(psi->weights, weights, psi->weights.size());
        if (biases) fill_cd(psi->biases, biases, psi->biases.size());
        if (final_weights) fill_cd(psi->final_weights, final_weights, psi->final_weights.size());
        psi->version++;
    });
}

int rbmhip_deep_set_translational_invariance(rbmhip_psi* psi, int on) {
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(psi->kind == 1, "not a PsiDeep");
        psi->translational_invariance = on ? 1 : 0;
    });
}

int rbmhip_psi_set_prefactor(rbmhip_psi* psi, double prefactor) {
    return guard([&] { check_psi(psi); psi->prefactor = prefactor; });
}

int rbmhip_psi_num_params(const rbmhip_psi* psi, int64_t* out) {
    return guard([&] { check_psi(psi); RBM_REQUIRE(out, "null output"); *out = psi->num_params(); });
}

int rbmhip_psi_destroy(rbmhip_psi* psi) {
    return guard([&] { delete psi; });
}

// ---- operators --------------------------------------------------------------
int rbmhip_op_create(int S, int L, const double* coefficients, const int32_t* types, const int32_t* indices,
                     rbmhip_op** out) {
    return guard([&] {
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(S >= 0 && L >= 0, "negative table shape");
        RBM_REQUIRE(S == 0 || (coefficients && types && indices && L > 0), "null operator tables");
        auto op = std::make_unique<rbmhip_op>();
        op->S = S; op->L = L;
        fill_cd(op->coeff, coefficients, (size_t)S);
        op->types.assign(types, types + (int64_t)S * L);
        op->idx.assign(indices, indices + (int64_t)S * L);
        canonicalise(op.get());
        *out = op.release();
    });
}

int rbmhip_op_info(const rbmhip_op* op, int* num_diagonal, int* num_flip_masks, int* hermitian, int* max_index) {
    return guard([&] {
        RBM_REQUIRE(op, "null operator");
        if (num_diagonal) *num_diagonal = (int)op->diag_coef.size();
        if (num_flip_masks) *num_flip_masks = (int)op->flip_begin.size() - 1;
        if (hermitian) *hermitian = op->hermitian ? 1 : 0;
        if (max_index) *max_index = op->max_index;
    });
}

int rbmhip_op_destroy(rbmhip_op* op) {
    return guard([&] { delete op; });
}

// ---- ensembles ----------------------------------------------------------------
int rbmhip_exact_create(int N, rbmhip_ens** out) {
    return guard([&] {
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(N >= 1 && N <= 40, "ExactSummation supports 1 <= N <= 40");
        auto e = std::make_unique<rbmhip_ens>();
        e->kind = 0; e->N = N; e->num_configs = (int64_t)1 << N;
        *out = e.release();
    });
}

int rbmhip_exact_set_total_z(rbmhip_ens* ens, int sector) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 0, "not an ExactSummation");
        RBM_REQUIRE(ens->N <= 32, "sector enumeration supports N <= 32");
        // ExactSummation.cu:26-45: keep basis indices whose total S_z
        // (2 popcount - N) equals the sector, in increasing order.
        ens->configs.clear();
        const uint64_t dim = (uint64_t)1 << ens->N;
        for (uint64_t idx = 0; idx < dim; idx++)
            if (2 * __builtin_popcountll(idx) - ens->N == sector) { ens->configs.push_back(idx); ens->configs.push_back(0); }
        ens->num_configs = (int64_t)(ens->configs.size() / 2);
        ens->has_sector = true;
        ens->sector = sector;
        for (auto& kv : ens->configs_dev) kv.second.release();
        ens->configs_dev.clear();
    });
}

int rbmhip_mc_create(uint32_t num_samples, uint32_t num_sweeps, uint32_t num_thermalization_sweeps,
                     uint32_t num_markov_chains, uint64_t seed, rbmhip_ens** out) {
    return guard([&] {
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(num_markov_chains > 0, "num_markov_chains must be > 0");
        RBM_REQUIRE(num_samples % num_markov_chains == 0,
                    "num_samples must be a multiple of num_markov_chains (MonteCarloLoop.cu:78)");
        auto e = std::make_unique<rbmhip_ens>();
        e->kind = 1;
        e->num_samples = num_samples; e->num_sweeps = num_sweeps; e->num_therm = num_thermalization_sweeps;
        e->num_chains = num_markov_chains; e->seed = seed;
        *out = e.release();
    });
}

int rbmhip_mc_copy(const rbmhip_ens* other, rbmhip_ens** out) {
    return guard([&] {
        RBM_REQUIRE(other && other->kind == 1 && out, "bad arguments");
        auto e = std::make_unique<rbmhip_ens>();
        e->kind = 1;
        e->num_samples = other->num_samples; e->num_sweeps = other->num_sweeps; e->num_therm = other->num_therm;
        e->num_chains = other->num_chains; e->seed = other->seed;
        e->has_sector = other->has_sector; e->sector = other->sector;
        e->fast_sweep = other->fast_sweep; e->fast_tries = other->fast_tries;
        e->persistent = other->persistent; e->record_angles = other->record_angles;
        e->resync_sweeps = other->resync_sweeps;
        *out = e.release();
    });
}

int rbmhip_mc_set_total_z(rbmhip_ens* ens, int sector) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 1, "set_total_z_symmetry applies to a MonteCarloLoop");
        ens->has_sector = true;
        ens->sector = sector;
    });
}

int rbmhip_mc_set_fast_sweep(rbmhip_ens* ens, uint32_t num_tries) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 1, "set_fast_sweep applies to a MonteCarloLoop");
        ens->fast_sweep = true;
        ens->fast_tries = num_tries;
    });
}

int rbmhip_mc_set_persistent(rbmhip_ens* ens, int on) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 1, "persistent chains are a MonteCarloLoop setting");
        ens->persistent = on != 0;
        ens->state_device = -1;   // the next call starts afresh either way
    });
}

int rbmhip_mc_get_persistent(const rbmhip_ens* ens, int* on) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(on, "null output");
        *on = ens->persistent ? 1 : 0;
    });
}

int rbmhip_mc_set_angle_resync(rbmhip_ens* ens, int sweeps) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 1, "angle re-synchronisation is a MonteCarloLoop setting");
        RBM_REQUIRE(sweeps >= 0 && sweeps <= (1 << 20), "sweeps must be in [0, 2^20]");
        ens->resync_sweeps = sweeps;
    });
}

int rbmhip_mc_record_angles(rbmhip_ens* ens, int on) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 1, "angle records are a MonteCarloLoop setting");
        ens->record_angles = on != 0;
        ens->angles_chains = 0;
    });
}

int rbmhip_mc_last_angles(const rbmhip_ens* ens, int64_t* chains, int* M, double* out) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(chains && M, "null output");
        *chains = ens->angles_chains;
        *M = ens->angles_M;
        if (!out || ens->angles_chains == 0) return;
        Ctx& c = ctx();
        const size_t n = (size_t)ens->angles_chains * ens->angles_M;
        if (ens->angles_precision == RBMHIP_FP64) {
            c.d2h(out, ens->angles.p, n * sizeof(cx<double>));
            c.sync();
        } else {
            std::vector<cx<float>> h(n);
            c.d2h(h.data(), ens->angles.p, n * sizeof(cx<float>));
            c.sync();
            for (size_t k = 0; k < n; k++) { out[2 * k] = h[k].re; out[2 * k + 1] = h[k].im; }
        }
    });
}

int rbmhip_ens_num_steps(const rbmhip_ens* ens, uint64_t* out) {
    return guard([&] { check_ens(ens); RBM_REQUIRE(out, "null output"); *out = (uint64_t)ens->num_steps(); });
}

int rbmhip_mc_acceptance(const rbmhip_ens* ens, uint64_t* acceptances, uint64_t* rejections) {
    return guard([&] {
        check_ens(ens);
        if (acceptances) *acceptances = ens->acc;
        if (rejections) *rejections = ens->rej;
    });
}

int rbmhip_mc_last_samples(const rbmhip_ens* ens, uint64_t* spins, double* log_psi, double* e_loc) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind != 0 && ens->rec_n > 0, "no Monte Carlo samples recorded yet");
        Ctx& c = ctx();
        RBM_REQUIRE(c.device == ens->rec_device, "samples were recorded on another device");
        if (ens->kind == 2) {
            // one record per run: expand to one row per sample
            const int64_t U = ens->run_n, n = ens->num_steps();
            std::vector<int64_t> run((size_t)n);
            std::vector<uint64_t> rs((size_t)U * 2);
            std::vector<cx<double>> rl((size_t)U), re((size_t)U);
            c.d2h(run.data(), ens->sample_run.p, run.size() * sizeof(int64_t));
            c.d2h(rs.data(), ens->spins.p, rs.size() * sizeof(uint64_t));
            c.d2h(rl.data(), ens->lp.p, rl.size() * sizeof(cx<double>));
            c.d2h(re.data(), ens->el.p, re.size() * sizeof(cx<double>));
            c.sync();
            for (int64_t k = 0; k < n; k++) {
                const int64_t r = run[(size_t)k];
                if (spins) { spins[2 * k] = rs[2 * r]; spins[2 * k + 1] = rs[2 * r + 1]; }
                if (log_psi) { log_psi[2 * k] = rl[r].re; log_psi[2 * k + 1] = rl[r].im; }
                if (e_loc) { e_loc[2 * k] = re[r].re; e_loc[2 * k + 1] = re[r].im; }
            }
            return;
        }
        if (spins) c.d2h(spins, ens->spins.p, (size_t)ens->rec_n * 2 * sizeof(uint64_t));
        if (log_psi) c.d2h(log_psi, ens->lp.p, (size_t)ens->rec_n * sizeof(cx<double>));
        if (e_loc) c.d2h(e_loc, ens->el.p, (size_t)ens->rec_n * sizeof(cx<double>));
        c.sync();
    });
}

int rbmhip_ens_destroy(rbmhip_ens* ens) {
    return guard([&] { delete ens; });
}

// ---- batched evaluation ---------------------------------------------------------
int rbmhip_log_psi(rbmhip_psi* psi, const uint64_t* spins, int64_t n, double* out) {
    const ApiRange range_("rbmhip_log_psi");
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(n >= 0 && (n == 0 || (spins && out)), "bad arguments");
        if (n == 0) return;
        Ctx& c = ctx();
        uint64_t* ds = reinterpret_cast<uint64_t*>(c.buf("in_spins", (size_t)n * 2 * sizeof(uint64_t)));
        cx<double>* lp = reinterpret_cast<cx<double>*>(c.buf("lp", (size_t)n * sizeof(cx<double>)));
        c.h2d(ds, spins, (size_t)n * 2 * sizeof(uint64_t));
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Records rec{nullptr, lp, nullptr, nullptr};
            eval_model<T>(c, psi, nullptr, n, ds, 0, rec);
        });
        c.d2h(out, lp, (size_t)n * sizeof(cx<double>));
        c.sync();
    });
}

int rbmhip_local_energy(rbmhip_psi* psi, rbmhip_op* op, const uint64_t* spins, int64_t n, double* out) {
    const ApiRange range_("rbmhip_local_energy");
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(op, "null operator");
        RBM_REQUIRE(n >= 0 && (n == 0 || (spins && out)), "bad arguments");
        check_op_fits(psi, op);
        if (n == 0) return;
        Ctx& c = ctx();
        uint64_t* ds = reinterpret_cast<uint64_t*>(c.buf("in_spins", (size_t)n * 2 * sizeof(uint64_t)));
        cx<double>* lp = reinterpret_cast<cx<double>*>(c.buf("lp", (size_t)n * sizeof(cx<double>)));
        cx<double>* el = reinterpret_cast<cx<double>*>(c.buf("el", (size_t)n * sizeof(cx<double>)));
        c.h2d(ds, spins, (size_t)n * 2 * sizeof(uint64_t));
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Records rec{nullptr, lp, el, nullptr};
            eval_model<T>(c, psi, op, n, ds, 0, rec);
        });
        c.d2h(out, el, (size_t)n * sizeof(cx<double>));
        c.sync();
    });
}

int rbmhip_local_energy_noalloc(rbmhip_psi* psi, rbmhip_op* op, const uint64_t* spins, int64_t n, double* out) {
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(op, "null operator");
        RBM_REQUIRE(n >= 0 && (n == 0 || (spins && out)), "bad arguments");
        RBM_REQUIRE(psi->kind == 0, "the allocation-free entry point evaluates RBM wavefunctions only");
        RBM_REQUIRE(psi->precision == RBMHIP_FP32 || psi->precision == RBMHIP_FP64, "unsupported precision");
        RBM_REQUIRE(n <= NOALLOC_MAX_CONFIGS, "n = " + std::to_string(n) + " exceeds the allocation-free limit of " +
                                                  std::to_string(NOALLOC_MAX_CONFIGS) + " configurations per call");
        check_op_fits(psi, op);
        if (n == 0) return;
        Ctx& c = ctx();
        if (!c.scratch)
            throw Error(RBMHIP_ERR_HIP, "the allocation-free scratch array is unavailable on this device: " + c.scratch_err);
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            local_energy_noalloc_impl<T>(c, psi, op, spins, n, out);
        });
    });
}

int rbmhip_o_k_vector(rbmhip_psi* psi, const uint64_t* spins, double* out) {
    const ApiRange range_("rbmhip_o_k_vector");
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(spins && out, "bad arguments");
        Ctx& c = ctx();
        const int64_t P = psi->num_params();
        uint64_t* ds = reinterpret_cast<uint64_t*>(c.buf("in_spins", 2 * sizeof(uint64_t)));
        c.h2d(ds, spins, 2 * sizeof(uint64_t));
        std::vector<cd> o((size_t)P, cd(0, 0));
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            const size_t K = psi->kind == 0 ? (size_t)psi->M : (size_t)(P - 2 * psi->N);
            void* t = c.buf("tanh", K * sizeof(cx<T>));
            cx<double>* lp = reinterpret_cast<cx<double>*>(c.buf("lp", sizeof(cx<double>)));
            Records rec{nullptr, lp, nullptr, t};
            eval_model<T>(c, psi, nullptr, 1, ds, 0, rec);
            std::vector<cx<T>> h(K);
            c.d2h(h.data(), t, K * sizeof(cx<T>));
            c.sync();
            if (psi->kind == 0) {
                // O_k = tanh(theta_j) for b_j, tanh(theta_j) s_i for W_ij, s_i for a_i
                const int N = psi->N, M = psi->M;
                uint64_t s0 = spins[0], s1 = spins[1];
                mask_words(s0, s1, N);
                for (int j = 0; j < M; j++) o[2 * N + j] = cd(h[j].re, h[j].im);
                for (int i = 0; i < N; i++) {
                    const double si = spin_of<double>(s0, s1, i);
                    for (int j = 0; j < M; j++) o[2 * N + M + (size_t)i * M + j] = si * cd(h[j].re, h[j].im);
                }
                if (psi->has_a)
                    for (int i = 0; i < N; i++) o[2 * N + M + (size_t)N * M + i] = spin_of<double>(s0, s1, i);
            } else {
                for (size_t k = 0; k < K; k++) o[2 * psi->N + k] = cd(h[k].re, h[k].im);
            }
        });
        std::memcpy(out, o.data(), sizeof(cd) * o.size());
    });
}

int rbmhip_psi_vector(rbmhip_psi* psi, double* out) {
    const ApiRange range_("rbmhip_psi_vector");
    return guard([&] {
        check_psi(psi);
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(psi->N <= 30, "psi_vector supports N <= 30");
        Ctx& c = ctx();
        const int64_t n = (int64_t)1 << psi->N;
        cx<double>* lp = reinterpret_cast<cx<double>*>(c.buf("lp", (size_t)n * sizeof(cx<double>)));
        cx<double>* v = reinterpret_cast<cx<double>*>(c.buf("vec", (size_t)n * sizeof(cx<double>)));
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Records rec{nullptr, lp, nullptr, nullptr};
            eval_model<T>(c, psi, nullptr, n, nullptr, 0, rec);
        });
        // PsiVector.cu:36-39: psi(s) = exp(log prefactor + log psi(s))
        c.launch("exp_vector", [&] { return launch_exp_vector(n, lp, std::log(psi->prefactor), v, c.stream); });
        c.d2h(out, v, (size_t)n * sizeof(cx<double>));
        c.sync();
    });
}

int rbmhip_psi_norm(rbmhip_psi* psi, rbmhip_ens* ens, double* out) {
    const ApiRange range_("rbmhip_psi_norm");
    return guard([&] {
        check_psi(psi);
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 0, "psi_norm needs an ExactSummation");
        RBM_REQUIRE(out, "null output");
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, nullptr, ens, false, false);
            int nb = (int)std::min<int64_t>(1024, (s.n + 255) / 256);
            double* part = reinterpret_cast<double*>(c.buf("lsepart", (size_t)nb * 2 * sizeof(double)));
            c.launch("lse", [&] { return launch_lse_partials(s.n, s.lp, std::log(psi->prefactor), nb, part, c.stream); });
            std::vector<double> h((size_t)nb * 2);
            c.d2h(h.data(), part, h.size() * sizeof(double));
            c.sync();
            double m = -INFINITY;
            for (int b = 0; b < nb; b++) m = std::max(m, h[2 * b]);
            double sum = 0.0;
            for (int b = 0; b < nb; b++)
                if (h[2 * b] != -INFINITY) sum += h[2 * b + 1] * std::exp(h[2 * b] - m);
            // sqrt(sum_s exp(x_s)) with x_s = 2 (log prefactor + Re log psi)
            *out = std::exp(0.5 * m) * std::sqrt(sum);
        });
    });
}

// ---- network functions ------------------------------------------------------------
int rbmhip_expectation(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* out) {
    const ApiRange range_("rbmhip_expectation");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(op && out, "bad arguments");
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, op, ens, false, false);
            Scalars r = energy_scalars<T>(c, psi, s, ens, nullptr);
            const double inv = 1.0 / (double)ens->num_steps();
            out[0] = r.v[1] * inv;
            out[1] = r.v[2] * inv;
        });
    });
}

int rbmhip_expectation_list(rbmhip_psi* psi, rbmhip_op* const* ops, int n_ops, rbmhip_ens* ens, double* out) {
    const ApiRange range_("rbmhip_expectation_list");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(n_ops >= 0 && (n_ops == 0 || (ops && out)), "bad arguments");
        for (int i = 0; i < n_ops; i++) { RBM_REQUIRE(ops[i], "null operator in list"); check_op_fits(psi, ops[i]); }
        if (n_ops == 0) return;
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            // one pass of the ensemble, then every operator on the same samples
            // (ExpectationValue.cu:448-518)
            Samples s = sample<T>(c, psi, nullptr, ens, true, false);
            const uint64_t* configs = ens->kind != 0 ? s.spins : ens_configs(c, ens);
            cx<double>* el = reinterpret_cast<cx<double>*>(c.buf("el_list", (size_t)s.n * sizeof(cx<double>)));
            cx<double>* lp2 = reinterpret_cast<cx<double>*>(c.buf("lp_list", (size_t)s.n * sizeof(cx<double>)));
            const double inv = 1.0 / (double)ens->num_steps();
            for (int i = 0; i < n_ops; i++) {
                Records rec{nullptr, lp2, el, nullptr};
                eval_model<T>(c, psi, ops[i], s.n, configs, 0, rec);
                Samples si = s;
                si.el = el;
                Scalars r = energy_scalars<T>(c, psi, si, ens, nullptr);
                out[2 * i] = r.v[1] * inv;
                out[2 * i + 1] = r.v[2] * inv;
            }
        });
    });
}

int rbmhip_fluctuation(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* fluctuation, double* expectation) {
    const ApiRange range_("rbmhip_fluctuation");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(op && fluctuation && expectation, "bad arguments");
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, op, ens, false, false);
            Scalars r = energy_scalars<T>(c, psi, s, ens, nullptr);
            const double inv = 1.0 / (double)ens->num_steps();
            const double Er = r.v[1] * inv, Ei = r.v[2] * inv, E2 = r.v[3] * inv;
            expectation[0] = Er;
            expectation[1] = Ei;
            *fluctuation = std::sqrt(E2 - (Er * Er + Ei * Ei));
        });
    });
}

}  // extern "C"

// sum_s w_s [(alpha E_s + beta) conj(O_k(s)) + (gamma E_s + delta) conj(E_k(s))]
// in parameter order, with the E_k term (Operator.hpp:145-179) summed over the
// virtual configurations s'_n of every sample in chunks of at most about 1 GiB
// of records.  Both contractions accumulate in FP64.  The reference's
// gradient (ExpectationValue.cu:243-253) is (1, -2<E>, 0, 1) / n.
static cx<double> cxd(cd z) { return mk(z.real(), z.imag()); }

template <typename T>
static std::vector<cd> gradient_ek_sums(Ctx& c, rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, const Samples& s,
                                        cd alpha, cd beta, cd gamma, cd delta) {
    const int64_t n = s.n;
    cx<double>* w = reinterpret_cast<cx<double>*>(c.buf("ek_w", (size_t)n * sizeof(cx<double>)));
    cx<double>* coef = reinterpret_cast<cx<double>*>(c.buf("coef", (size_t)n * sizeof(cx<double>)));
    CoefArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = n; a.exact = ens->kind == 0; a.num_configs = (double)ens->num_configs; a.mult = s.mult;
    a.log_prefactor = std::log(psi->prefactor); a.lp = s.lp; a.el = s.el;
    a.mode = 4; a.coef = w;
    c.launch("coef", [&] { return launch_coef(a, c.stream); });
    a.mode = 5; a.coef = coef; a.alpha = cxd(alpha); a.e_mean = cxd(beta);
    c.launch("coef", [&] { return launch_coef(a, c.stream); });
    std::vector<cd> R = contract_params<T>(c, psi, n, s.tanh, s.spins, coef, true);
    const OpDev od = op_dev(c, op);
    const int T1 = od.n_terms + 1;
    const size_t per_v = 4 * sizeof(cx<double>) + (psi->kind == 0 ? (size_t)psi->M * sizeof(cx<T>) : 0);
    const int64_t cmax = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)((1ull << 30) / (per_v * (size_t)T1))));
    const int64_t vmax = cmax * T1;
    uint64_t* vspins = reinterpret_cast<uint64_t*>(c.buf("ek_spins", (size_t)vmax * 2 * sizeof(uint64_t)));
    cx<double>* vfac = reinterpret_cast<cx<double>*>(c.buf("ek_fac", (size_t)vmax * sizeof(cx<double>)));
    cx<double>* vlp = reinterpret_cast<cx<double>*>(c.buf("ek_lp", (size_t)vmax * sizeof(cx<double>)));
    cx<double>* vcoef = reinterpret_cast<cx<double>*>(c.buf("ek_coef", (size_t)vmax * sizeof(cx<double>)));
    void* vtanh = psi->kind == 0 ? c.buf("ek_tanh", (size_t)vmax * psi->M * sizeof(cx<T>)) : nullptr;
    for (int64_t i0 = 0; i0 < n; i0 += cmax) {
        const int64_t cnt = std::min<int64_t>(cmax, n - i0), V = cnt * T1;
        c.launch("ek_expand", [&] { return launch_ek_expand(od, i0, cnt, s.spins, vspins, vfac, c.stream); });
        Records rec{nullptr, vlp, nullptr, vtanh};
        eval_model<T>(c, psi, nullptr, V, vspins, 0, rec, "ek_eval");
        c.launch("ek_coef", [&] {
            return launch_ek_coef(i0, cnt, T1, w, s.lp, gamma == cd(0, 0) ? nullptr : s.el, cxd(gamma), cxd(delta), vlp,
                                  vfac, vcoef, c.stream);
        });
        const std::vector<cd> Rv = contract_params<T>(c, psi, V, vtanh, vspins, vcoef, true);
        for (size_t k = 0; k < R.size(); k++) R[k] += Rv[k];
    }
    return R;
}

static int gradient_impl(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, int estimator, double* grad, double* expectation) {
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(op && grad && expectation, "bad arguments");
        RBM_REQUIRE(estimator >= RBMHIP_GRAD_AUTO && estimator <= RBMHIP_GRAD_HERMITIAN, "unknown gradient estimator");
        if (estimator == RBMHIP_GRAD_HERMITIAN)
            RBM_REQUIRE(op->hermitian, "the Hermitian gradient estimator needs a Hermitian operator");
        const bool use_ek = estimator == RBMHIP_GRAD_EK || (estimator == RBMHIP_GRAD_AUTO && !op->hermitian);
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, op, ens, true, true);
            Scalars r = energy_scalars<T>(c, psi, s, ens, nullptr);
            const int64_t P = psi->num_params();
            const double n = (double)ens->num_steps();
            const cd E(r.v[1] / n, r.v[2] / n);
            if (use_ek) {
                const std::vector<cd> R = gradient_ek_sums<T>(c, psi, op, ens, s, cd(1, 0), -2.0 * E, cd(0, 0), cd(1, 0));
                for (int64_t k = 0; k < P; k++) {
                    const cd g = R[(size_t)k] / n;
                    grad[2 * k] = g.real();
                    grad[2 * k + 1] = g.imag();
                }
                expectation[0] = E.real();
                expectation[1] = E.imag();
                return;
            }
            // ExpectationValue.cu:243-253 with conj<E_k> = <E conj(O_k)> (Hermitian):
            // g_k = 2 <E conj O_k> - 2 <E> conj<O_k> = (2/n) sum_s w (E - <E>) conj(O_k)
            cx<double>* coef = reinterpret_cast<cx<double>*>(c.buf("coef", (size_t)s.n * sizeof(cx<double>)));
            CoefArgs a;
            std::memset(&a, 0, sizeof(a));
            a.n = s.n; a.mode = 2; a.exact = ens->kind == 0; a.num_configs = (double)ens->num_configs; a.mult = s.mult;
            a.log_prefactor = std::log(psi->prefactor); a.lp = s.lp; a.el = s.el; a.coef = coef;
            a.e_mean = mk(E.real(), E.imag());
            c.launch("coef", [&] { return launch_coef(a, c.stream); });
            std::vector<cd> R = contract_params<T>(c, psi, s.n, s.tanh, s.spins, coef);
            for (int64_t k = 0; k < P; k++) {
                const cd g = (2.0 / n) * R[(size_t)k];
                grad[2 * k] = g.real();
                grad[2 * k + 1] = g.imag();
            }
            expectation[0] = E.real();
            expectation[1] = E.imag();
        });
    });
}

extern "C" {

int rbmhip_gradient(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* grad, double* expectation) {
    const ApiRange range_("rbmhip_gradient");
    return gradient_impl(psi, op, ens, RBMHIP_GRAD_AUTO, grad, expectation);
}

int rbmhip_gradient_estimator(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, int estimator, double* grad,
                              double* expectation) {
    const ApiRange range_("rbmhip_gradient_estimator");
    return gradient_impl(psi, op, ens, estimator, grad, expectation);
}

int rbmhip_fluctuation_gradient(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double* grad) {
    const ApiRange range_("rbmhip_fluctuation_gradient");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(op && grad, "bad arguments");
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, op, ens, true, true);
            Scalars r = energy_scalars<T>(c, psi, s, ens, nullptr);
            const int64_t P = psi->num_params();
            const double n = (double)ens->num_steps();
            const cd E(r.v[1] / n, r.v[2] / n);
            const double E2 = r.v[3] / n, nE = std::norm(E);
            const double fl = std::sqrt(E2 - nE);
            // ExpectationValue.cu:322-330:
            // [2 conj(<E* E_k> - <E>* <E_k>) - 2 <E>* <E O*> + 2 conj<O> (2|<E>|^2 - <|E|^2>)] / (2 sigma)
            const std::vector<cd> R = gradient_ek_sums<T>(c, psi, op, ens, s, -2.0 * std::conj(E), cd(2.0 * (2.0 * nE - E2), 0),
                                                          cd(2, 0), -2.0 * E);
            for (int64_t k = 0; k < P; k++) {
                const cd g = R[(size_t)k] / (n * 2.0 * fl);
                grad[2 * k] = g.real();
                grad[2 * k + 1] = g.imag();
            }
        });
    });
}

int rbmhip_corrected(rbmhip_psi* psi, rbmhip_op* op, rbmhip_op* op2, rbmhip_ens* ens, double* value, double* norm) {
    const ApiRange range_("rbmhip_corrected");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(op && value && norm, "bad arguments");
        (void)op2;   // the reference evaluates both local energies with operator_ (ExpectationValue.cu:102-106)
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, op, ens, false, false);
            double* sc = reinterpret_cast<double*>(c.buf("sc", (size_t)s.n * SC_STRIDE * sizeof(double)));
            CoefArgs a;
            std::memset(&a, 0, sizeof(a));
            a.n = s.n; a.mode = 6; a.exact = ens->kind == 0; a.num_configs = (double)ens->num_configs; a.mult = s.mult;
            a.log_prefactor = std::log(psi->prefactor); a.lp = s.lp; a.el = s.el; a.sc = sc;
            c.launch("coef", [&] { return launch_coef(a, c.stream); });
            Scalars r = reduce_scalars(c, s.n, sc);
            const double n = (double)ens->num_steps();
            const cd A(r.v[0] / n, r.v[1] / n);
            const double X = r.v[2] / n;
            const cd v = A / X;
            value[0] = v.real();
            value[1] = v.imag();
            *norm = X;
        });
    });
}

int rbmhip_difference(rbmhip_psi* psi, rbmhip_psi* psi_prime, rbmhip_op* const* ops, int n_ops, rbmhip_ens* ens,
                      double* out) {
    const ApiRange range_("rbmhip_difference");
    return guard([&] {
        check_psi(psi); check_psi(psi_prime); check_ens(ens);
        RBM_REQUIRE(ops && n_ops >= 1 && out, "bad arguments");
        RBM_REQUIRE(psi->N == psi_prime->N, "psi and psi_prime have different N");
        for (int i = 0; i < n_ops; i++) {
            RBM_REQUIRE(ops[i] != nullptr, "null operator");
            check_op_fits(psi, ops[i]);
        }
        Ctx& c = ctx();
        Samples s;
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            s = sample<T>(c, psi, nullptr, ens, true, false);
        });
        const uint64_t* configs = s.spins;
        cx<double>* lp2 = reinterpret_cast<cx<double>*>(c.buf("lp2", (size_t)s.n * sizeof(cx<double>)));
        cx<double>* el = reinterpret_cast<cx<double>*>(c.buf("diff_el", (size_t)s.n * sizeof(cx<double>)));
        cx<double>* el2 = reinterpret_cast<cx<double>*>(c.buf("diff_el2", (size_t)s.n * sizeof(cx<double>)));
        double* sc = reinterpret_cast<double*>(c.buf("sc", (size_t)s.n * SC_STRIDE * sizeof(double)));
        with_precision(psi_prime->precision, [&](auto tag) {
            using T = decltype(tag);
            Records rec{nullptr, lp2, nullptr, nullptr};
            eval_model<T>(c, psi_prime, nullptr, s.n, configs, 0, rec);
        });
        CoefArgs a;
        std::memset(&a, 0, sizeof(a));
        a.n = s.n; a.exact = ens->kind == 0; a.num_configs = (double)ens->num_configs; a.mult = s.mult;
        a.log_prefactor = std::log(psi->prefactor); a.lp = s.lp; a.lp2 = lp2; a.sc = sc;
        a.mode = 8;
        c.launch("coef", [&] { return launch_coef(a, c.stream); });
        const double n = (double)ens->num_steps();
        const double ravg = reduce_scalars(c, s.n, sc).v[0] / n;
        for (int i = 0; i < n_ops; i++) {
            with_precision(psi->precision, [&](auto tag) {
                using T = decltype(tag);
                Records rec{nullptr, nullptr, el, nullptr};
                eval_model<T>(c, psi, ops[i], s.n, configs, 0, rec);
            });
            with_precision(psi_prime->precision, [&](auto tag) {
                using T = decltype(tag);
                Records rec{nullptr, nullptr, el2, nullptr};
                eval_model<T>(c, psi_prime, ops[i], s.n, configs, 0, rec);
            });
            a.mode = 7; a.el = el; a.el2 = el2;
            c.launch("coef", [&] { return launch_coef(a, c.stream); });
            Scalars r = reduce_scalars(c, s.n, sc);
            const cd av(r.v[0] / n, r.v[1] / n), bv(r.v[2] / n, r.v[3] / n);
            const cd d = av - bv / ravg;
            out[2 * i] = d.real();
            out[2 * i + 1] = d.imag();
        }
    });
}

int rbmhip_s_matrix(rbmhip_psi* psi, rbmhip_ens* ens, double* out) {
    const ApiRange range_("rbmhip_s_matrix");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(psi->kind == 0, "get_S_matrix is bound for Psi only (main.cpp:557-560)");
        Ctx& c = ctx();
        const int64_t P = psi->num_params();
        RBM_REQUIRE(P <= 8192, "S matrix larger than 8192 x 8192");
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, nullptr, ens, true, true);
            cx<double>* S = reinterpret_cast<cx<double>*>(c.buf("smat", (size_t)(P * P) * sizeof(cx<double>)));
            const double inv = 1.0 / (double)ens->num_steps();
            c.launch("s_matrix", [&] { return launch_s_matrix_ref<T>(s.n, psi->N, psi->M, psi->has_a, s.tanh, s.spins, inv, S, c.stream); });
            c.d2h(out, S, (size_t)(P * P) * sizeof(cx<double>));
            c.sync();
        });
    });
}

int rbmhip_s_matrix_centred(rbmhip_psi* psi, rbmhip_ens* ens, double* out) {
    const ApiRange range_("rbmhip_s_matrix_centred");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(out, "null output");
        Ctx& c = ctx();
        const int64_t P = psi->num_params() - 2 * (int64_t)psi->N;
        RBM_REQUIRE(P <= 16384, "S matrix larger than 16384 x 16384");
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, nullptr, ens, true, true);
            cx<double>* S = s_matrix_centred<T>(c, psi, ens, s, 0.0);
            c.d2h(out, S, (size_t)(P * P) * sizeof(cx<double>));
            c.sync();
        });
    });
}

int rbmhip_cholesky_solve(int n, const double* A, const double* b, double* x) {
    const ApiRange range_("rbmhip_cholesky_solve");
    return guard([&] {
        RBM_REQUIRE(n > 0 && A && b && x, "bad arguments");
        Ctx& c = ctx();
        cx<double>* dA = reinterpret_cast<cx<double>*>(c.buf("chol_A", (size_t)n * n * sizeof(cx<double>)));
        cx<double>* dx = reinterpret_cast<cx<double>*>(c.buf("chol_x", (size_t)n * sizeof(cx<double>)));
        c.h2d(dA, A, (size_t)n * n * sizeof(cx<double>));
        c.h2d(dx, b, (size_t)n * sizeof(cx<double>));
        cholesky_solve_device(c, n, dA, dx);
        c.d2h(x, dx, (size_t)n * sizeof(cx<double>));
        c.sync();
    });
}

int rbmhip_sr_update(rbmhip_psi* psi, rbmhip_op* op, rbmhip_ens* ens, double eps, double* update, double* expectation) {
    const ApiRange range_("rbmhip_sr_update");
    return guard([&] {
        check_psi(psi); check_ens(ens);
        RBM_REQUIRE(op && update && expectation, "bad arguments");
        RBM_REQUIRE(eps >= 0.0, "eps must be >= 0");
        if (!op->hermitian)
            throw Error(RBMHIP_ERR_UNSUPPORTED, "stochastic reconfiguration needs a Hermitian operator");
        Ctx& c = ctx();
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            Samples s = sample<T>(c, psi, op, ens, true, true);
            Scalars r = energy_scalars<T>(c, psi, s, ens, nullptr);
            const int64_t Pall = psi->num_params();
            const int N = psi->N;
            const int P = (int)(Pall - 2 * (int64_t)N);
            const double n = (double)ens->num_steps();
            const cd E(r.v[1] / n, r.v[2] / n);
            cx<double>* coef = reinterpret_cast<cx<double>*>(c.buf("coef", (size_t)s.n * sizeof(cx<double>)));
            CoefArgs a;
            std::memset(&a, 0, sizeof(a));
            a.n = s.n; a.mode = 2; a.exact = ens->kind == 0; a.num_configs = (double)ens->num_configs; a.mult = s.mult;
            a.log_prefactor = std::log(psi->prefactor); a.lp = s.lp; a.el = s.el; a.coef = coef;
            a.e_mean = mk(E.real(), E.imag());
            c.launch("coef", [&] { return launch_coef(a, c.stream); });
            const std::vector<cd> R = contract_params<T>(c, psi, s.n, s.tanh, s.spins, coef);
            std::vector<cd> g((size_t)P);
            for (int k = 0; k < P; k++) g[(size_t)k] = (2.0 / n) * R[(size_t)(2 * N + k)];
            cx<double>* S = s_matrix_centred<T>(c, psi, ens, s, eps);
            cx<double>* dx = reinterpret_cast<cx<double>*>(c.buf("sr_x", (size_t)P * sizeof(cx<double>)));
            c.h2d(dx, g.data(), g.size() * sizeof(cx<double>));
            cholesky_solve_device(c, P, S, dx);
            std::vector<cd> x((size_t)P);
            c.d2h(x.data(), dx, x.size() * sizeof(cx<double>));
            c.sync();
            for (int64_t k = 0; k < Pall; k++) {
                const cd v = k < 2 * N ? cd(0, 0) : x[(size_t)(k - 2 * N)];
                update[2 * k] = v.real();
                update[2 * k + 1] = v.imag();
            }
            expectation[0] = E.real();
            expectation[1] = E.imag();
        });
    });
}

// HilbertSpaceDistance::distance / gradient; axis (4 N host values: alpha' -
// alpha, beta' - beta, cos and sin of alpha' + alpha) selects the free quantum
// axis form, HilbertSpaceDistance.cu:47-80 and 109-122.
static int hs_distance_impl(rbmhip_psi* psi, rbmhip_psi* psi_prime, rbmhip_op* op, int is_unitary, rbmhip_ens* ens,
                            int want_grad, double nu, const std::vector<double>* axis, double* distance, double* grad) {
    return guard([&] {
        check_psi(psi); check_psi(psi_prime); check_ens(ens);
        RBM_REQUIRE(op && distance && (!want_grad || grad), "bad arguments");
        RBM_REQUIRE(psi->N == psi_prime->N, "psi and psi_prime have different N");
        Ctx& c = ctx();
        Samples s;
        with_precision(psi->precision, [&](auto tag) {
            using T = decltype(tag);
            s = sample<T>(c, psi, op, ens, true, false);
        });
        const uint64_t* configs = ens->kind != 0 ? s.spins : (ens->has_sector ? ens_configs(c, ens) : nullptr);
        // HilbertSpaceDistance.cu:40-44: log psi'(s) on the same samples (and tanh for O_k(psi'))
        cx<double>* lp2 = reinterpret_cast<cx<double>*>(c.buf("lp2", (size_t)s.n * sizeof(cx<double>)));
        void* tanh2 = nullptr;
        uint64_t* spins2 = nullptr;
        with_precision(psi_prime->precision, [&](auto tag) {
            using T = decltype(tag);
            if (want_grad) {
                if (psi_prime->kind == 0) tanh2 = c.buf("tanh2", (size_t)s.n * (size_t)psi_prime->M * sizeof(cx<T>));
                if (ens->kind == 0 && !s.spins) spins2 = reinterpret_cast<uint64_t*>(c.buf("spins", (size_t)s.n * 2 * sizeof(uint64_t)));
            }
            Records rec{spins2, lp2, nullptr, tanh2};
            eval_model<T>(c, psi_prime, nullptr, s.n, configs, 0, rec);
        });
        if (psi->kind == 0 && psi_prime->kind == 0 && psi->M == psi_prime->M && psi->precision == psi_prime->precision) {
            // nearby RBMs: log psi' - log psi from the angle difference, unit by
            // unit, instead of the difference of two sums of M log-cosh values
            with_precision(psi->precision, [&](auto tag) {
                using T = decltype(tag);
                const RbmDev<T> md = rbm_dev<T>(c, psi);
                const int Nn = psi->N, Mm = psi->M, ld = md.ld;
                std::vector<cx<T>> hw((size_t)Nn * ld, mk<T>(0, 0)), hb((size_t)ld, mk<T>(0, 0));
                for (int i = 0; i < Nn; i++)
                    for (int j = 0; j < Mm; j++) {
                        const cd d = psi_prime->W[(size_t)i * Mm + j] - psi->W[(size_t)i * Mm + j];
                        hw[(size_t)i * ld + j] = mk<T>((T)d.real(), (T)d.imag());
                    }
                for (int j = 0; j < Mm; j++) {
                    const cd d = psi_prime->b[j] - psi->b[j];
                    hb[j] = mk<T>((T)d.real(), (T)d.imag());
                }
                void* dW = c.buf("hs_dW", hw.size() * sizeof(cx<T>));
                void* db = c.buf("hs_db", hb.size() * sizeof(cx<T>));
                RBM_HIPCHK(hipMemcpy(dW, hw.data(), hw.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
                RBM_HIPCHK(hipMemcpy(db, hb.data(), hb.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
                cx<double>* da = nullptr;
                if (psi->has_a || psi_prime->has_a) {
                    std::vector<cx<double>> ha((size_t)Nn);
                    for (int i = 0; i < Nn; i++) {
                        const cd d = (psi_prime->has_a ? psi_prime->a[i] : cd(0, 0)) - (psi->has_a ? psi->a[i] : cd(0, 0));
                        ha[i] = mk(d.real(), d.imag());
                    }
                    da = reinterpret_cast<cx<double>*>(c.buf("hs_da", ha.size() * sizeof(cx<double>)));
                    RBM_HIPCHK(hipMemcpy(da, ha.data(), ha.size() * sizeof(cx<double>), hipMemcpyHostToDevice));
                }
                c.launch("hs_log_ratio", [&] {
                    return launch_rbm_log_ratio<T>(md, dW, db, da, s.n, configs, 0, s.lp, lp2, c.num_cus * 8, c.stream);
                });
            });
        }
        const uint64_t* spins_for_grad = s.spins ? s.spins : spins2;
        // free quantum axis: log psi' on the N single-flip neighbours of every
        // sample (chunks of at most 2^26 configurations), ratios r_i and the
        // first-order axis term added to log psi'
        const int N = psi->N;
        double* axis_dev = nullptr;
        cx<double>* flip_r = nullptr;
        if (axis) {
            RBM_REQUIRE((double)s.n * N * sizeof(cx<double>) <= (double)(32ll << 30),
                        "free_quantum_axis keeps N flip ratios per sample; this ensemble needs more than 32 GiB for them");
            axis_dev = reinterpret_cast<double*>(c.buf("axis", axis->size() * sizeof(double)));
            c.h2d(axis_dev, axis->data(), axis->size() * sizeof(double));
            flip_r = reinterpret_cast<cx<double>*>(c.buf("axis_ratio", (size_t)s.n * N * sizeof(cx<double>)));
            const int per = (int)std::max<int64_t>(1, std::min<int64_t>(N, (1ll << 26) / s.n));
            uint64_t* fc = reinterpret_cast<uint64_t*>(c.buf("axis_configs", (size_t)s.n * per * 2 * sizeof(uint64_t)));
            for (int i0 = 0; i0 < N; i0 += per) {
                const int cnt = std::min(per, N - i0);
                c.launch("axis_flip", [&] { return launch_flip_configs(s.n, configs, 0, i0, cnt, fc, c.stream); });
                with_precision(psi_prime->precision, [&](auto tag) {
                    using T = decltype(tag);
                    Records rec{nullptr, flip_r + (size_t)i0 * s.n, nullptr, nullptr};
                    eval_model<T>(c, psi_prime, nullptr, s.n * cnt, fc, 0, rec, "axis_eval");
                });
            }
            c.launch("axis_log_psi", [&] { return launch_axis_log_psi(s.n, N, configs, 0, axis_dev, lp2, flip_r, c.stream); });
        }
        double* sc = reinterpret_cast<double*>(c.buf("sc", (size_t)s.n * SC_STRIDE * sizeof(double)));
        CoefArgs a;
        std::memset(&a, 0, sizeof(a));
        a.n = s.n; a.mode = 1; a.exact = ens->kind == 0; a.is_unitary = is_unitary ? 1 : 0; a.mult = s.mult;
        a.num_configs = (double)ens->num_configs; a.log_prefactor = std::log(psi->prefactor);
        a.lp = s.lp; a.el = s.el; a.lp2 = lp2; a.sc = sc;
        c.launch("coef", [&] { return launch_coef(a, c.stream); });
        Scalars r = reduce_scalars(c, s.n, sc);
        const double n = (double)ens->num_steps();
        const cd omega(r.v[0] / n, r.v[1] / n);
        const double ratio = r.v[2] / n, nrm = r.v[3] / n;
        // HilbertSpaceDistance.cu:216-220, 248-262
        const double u = std::norm(omega);
        const double v = nrm * ratio;
        const double d = std::sqrt(1.0 - u / v);
        *distance = d;
        if (want_grad) {
            // g_k = -(conj(w) <w O*> v - u n <r O*>) / v^2 / d^nu
            //     = -conj(w) n_avg / (n v^2 d^nu) * sum_s w (r_avg omega - omega_avg r) conj(O_k)
            cx<double>* coef = reinterpret_cast<cx<double>*>(c.buf("coef", (size_t)s.n * sizeof(cx<double>)));
            a.mode = 3; a.coef = coef; a.sc = nullptr;
            a.omega_mean = mk(omega.real(), omega.imag());
            a.ratio_mean = ratio;
            c.launch("coef", [&] { return launch_coef(a, c.stream); });
            std::vector<cd> R;
            with_precision(psi_prime->precision, [&](auto tag) {
                using T = decltype(tag);
                R = contract_params<T>(c, psi_prime, s.n, tanh2, spins_for_grad, coef);
            });
            if (axis) {
                const int slices = (int)std::max<int64_t>(1, std::min<int64_t>(256, s.n / 4096));
                cx<double>* part = reinterpret_cast<cx<double>*>(c.buf("axis_part", (size_t)slices * 2 * N * sizeof(cx<double>)));
                cx<double>* out = reinterpret_cast<cx<double>*>(c.buf("axis_out", (size_t)2 * N * sizeof(cx<double>)));
                c.launch("axis_contract", [&] {
                    return launch_axis_contract(s.n, N, configs, 0, axis_dev, flip_r, coef, slices, part, c.stream);
                });
                c.launch("partials_sum", [&] { return launch_partials_sum(slices, 2 * N, part, out, c.stream); });
                std::vector<cd> h((size_t)2 * N);
                c.d2h(h.data(), out, h.size() * sizeof(cx<double>));
                c.sync();
                for (int k = 0; k < 2 * N; k++) R[(size_t)k] = h[(size_t)k];
            }
            const int64_t P = psi_prime->num_params();
            const double pref = std::pow(d, (double)(float)nu);
            const cd fac = -std::conj(omega) * nrm / (n * v * v * pref);
            for (int64_t k = 0; k < P; k++) {
                const cd g = fac * R[(size_t)k];
                grad[2 * k] = g.real();
                grad[2 * k + 1] = g.imag();
            }
        }
    });
}

int rbmhip_hs_distance(rbmhip_psi* psi, rbmhip_psi* psi_prime, rbmhip_op* op, int is_unitary, rbmhip_ens* ens,
                       int want_grad, double nu, double* distance, double* grad) {
    const ApiRange range_("rbmhip_hs_distance");
    return hs_distance_impl(psi, psi_prime, op, is_unitary, ens, want_grad, nu, nullptr, distance, grad);
}

int rbmhip_hs_distance_axis(rbmhip_psi* psi, rbmhip_psi* psi_prime, rbmhip_op* op, int is_unitary, rbmhip_ens* ens,
                            int want_grad, double nu, const double* alpha, const double* beta, const double* alpha_prime,
                            const double* beta_prime, double* distance, double* grad) {
    const ApiRange range_("rbmhip_hs_distance_axis");
    int rc = guard([&] {
        check_psi(psi);
        RBM_REQUIRE(alpha && beta && alpha_prime && beta_prime, "null quantum-axis angles");
    });
    if (rc != 0) return rc;
    const int N = psi->N;
    std::vector<double> axis((size_t)4 * N);
    for (int i = 0; i < N; i++) {
        axis[(size_t)i] = alpha_prime[i] - alpha[i];
        axis[(size_t)(N + i)] = beta_prime[i] - beta[i];
        axis[(size_t)(2 * N + i)] = std::cos(alpha_prime[i] + alpha[i]);
        axis[(size_t)(3 * N + i)] = std::sin(alpha_prime[i] + alpha[i]);
    }
    return hs_distance_impl(psi, psi_prime, op, is_unitary, ens, want_grad, nu, &axis, distance, grad);
}

// ---- timing -------------------------------------------------------------------------
int rbmhip_synchronize(void) {
    return guard([&] { ctx().sync(); });
}

int rbmhip_timing_enable(int on) {
    return guard([&] { Ctx& c = ctx(); c.sync(); c.timing = on != 0; });
}

int rbmhip_timing_reset(void) {
    return guard([&] { Ctx& c = ctx(); c.sync(); c.totals.clear(); });
}

int rbmhip_timing_query(const char* name, double* total_ms, int64_t* launches) {
    return guard([&] {
        RBM_REQUIRE(name, "null name");
        Ctx& c = ctx();
        c.sync();
        auto it = c.totals.find(name);
        if (total_ms) *total_ms = it == c.totals.end() ? 0.0 : it->second.first;
        if (launches) *launches = it == c.totals.end() ? 0 : it->second.second;
    });
}

// ---- masked autoregressive proposal ---------------------------------------------
int rbmhip_made_create(int N, int H, const float* params, uint64_t seed, rbmhip_made** out) {
    return guard([&] {
        RBM_REQUIRE(out, "null output");
        RBM_REQUIRE(N >= 1 && N <= 128, "the proposal network needs 1 <= N <= 128");
        RBM_REQUIRE(H >= 1 && made_pad(H) <= 512, "the proposal network needs 1 <= hidden units <= 512");
        Ctx& c = ctx();
        auto md = std::make_unique<rbmhip_made>();
        md->s.N = N; md->s.H = H; md->s.Np = made_pad(N); md->s.Hp = made_pad(H);
        RBM_REQUIRE((size_t)2 * N * md->s.Hp * sizeof(uint16_t) <= 160 * 1024,
                    "the bfloat16 weights of this network do not fit in one workgroup's LDS (N * hidden too large)");
        const int Np = md->s.Np, Hp = md->s.Hp;
        md->off_b1 = (int64_t)Hp * Np;
        md->off_W2 = md->off_b1 + Hp;
        md->off_b2 = md->off_W2 + (int64_t)Np * Hp;
        md->padded = md->off_b2 + Np;
        md->device = c.device;
        const std::vector<float> flat = params ? std::vector<float>(params, params + md->num_params())
                                               : made_init_params(N, H, seed);
        const std::vector<float> dev = made_pad_params(md.get(), flat.data());
        const size_t bytes = (size_t)md->padded * sizeof(float);
        md->params.ensure(bytes); md->grad.ensure(bytes); md->m.ensure(bytes); md->v.ensure(bytes);
        c.h2d(md->params.p, dev.data(), bytes);
        RBM_HIPCHK(launch_zero(md->padded, reinterpret_cast<float*>(md->grad.p), c.stream));
        RBM_HIPCHK(launch_zero(md->padded, reinterpret_cast<float*>(md->m.p), c.stream));
        RBM_HIPCHK(launch_zero(md->padded, reinterpret_cast<float*>(md->v.p), c.stream));
        c.sync();
        *out = md.release();
    });
}

int rbmhip_made_num_params(const rbmhip_made* made, int64_t* out) {
    return guard([&] { RBM_REQUIRE(made && out, "bad arguments"); *out = made->num_params(); });
}

int rbmhip_made_get_params(rbmhip_made* made, float* out) {
    return guard([&] {
        RBM_REQUIRE(made && out, "bad arguments");
        Ctx& c = ctx();
        made_check_device(c, made);
        std::vector<float> dev((size_t)made->padded);
        c.d2h(dev.data(), made->params.p, dev.size() * sizeof(float));
        c.sync();
        const std::vector<float> flat = made_unpad_params(made, dev);
        std::memcpy(out, flat.data(), flat.size() * sizeof(float));
    });
}

int rbmhip_made_set_params(rbmhip_made* made, const float* params) {
    return guard([&] {
        RBM_REQUIRE(made && params, "bad arguments");
        Ctx& c = ctx();
        made_check_device(c, made);
        const std::vector<float> dev = made_pad_params(made, params);
        c.h2d(made->params.p, dev.data(), dev.size() * sizeof(float));
        c.sync();
        made->inf_valid = false;
    });
}

int rbmhip_made_check_launches(int N, int H, int64_t n, int batch) {
    return guard([&] {
        RBM_REQUIRE(N >= 1 && N <= 128 && H >= 1 && made_pad(H) <= 512 && n > 0 && batch > 0 && batch <= (1 << 24),
                    "bad arguments");
        made_check_extents(made_plan(N, H, n, batch));
    });
}

int rbmhip_made_reset_optimizer(rbmhip_made* made) {
    return guard([&] {
        RBM_REQUIRE(made, "null handle");
        Ctx& c = ctx();
        made_check_device(c, made);
        RBM_HIPCHK(launch_zero(made->padded, reinterpret_cast<float*>(made->m.p), c.stream));
        RBM_HIPCHK(launch_zero(made->padded, reinterpret_cast<float*>(made->v.p), c.stream));
        c.sync();
        made->t = 0;
    });
}

int rbmhip_made_loss_grad(rbmhip_made* made, const uint64_t* spins, int64_t n, double* loss, float* grad) {
    const ApiRange range_("rbmhip_made_loss_grad");
    return guard([&] {
        RBM_REQUIRE(made && spins && n > 0, "bad arguments");
        RBM_REQUIRE(n <= (1 << 24), "batch too large");
        Ctx& c = ctx();
        made_check_device(c, made);
        const int B = (int)n, Bp = made_pad(B);
        made_check_extents(made_plan(made->s.N, made->s.H, n, B));
        uint64_t* ds = reinterpret_cast<uint64_t*>(c.buf("made_in", (size_t)n * 2 * sizeof(uint64_t)));
        c.h2d(ds, spins, (size_t)n * 2 * sizeof(uint64_t));
        float* X = reinterpret_cast<float*>(c.buf("made_x", (size_t)Bp * made->s.Np * sizeof(float)));
        c.launch("made_gather", [&] { return launch_made_gather(Bp, B, made->s.N, made->s.Np, ds, nullptr, X, c.stream); });
        const int tiles = made_loss_tiles(made, Bp);
        double* part = reinterpret_cast<double*>(c.buf("made_part", (size_t)tiles * sizeof(double)));
        made_forward_backward(c, made, B, Bp, X, part);
        std::vector<double> hp((size_t)tiles);
        std::vector<float> dev((size_t)made->padded);
        c.d2h(hp.data(), part, hp.size() * sizeof(double));
        c.d2h(dev.data(), made->grad.p, dev.size() * sizeof(float));
        c.sync();
        double sum = 0.0;
        for (double v : hp) sum += v;
        if (loss) *loss = -sum / (double)B;
        if (grad) {
            const std::vector<float> flat = made_unpad_params(made, dev);
            std::memcpy(grad, flat.data(), flat.size() * sizeof(float));
        }
    });
}

int rbmhip_made_train(rbmhip_made* made, const uint64_t* spins, int64_t n, int steps, int batch, double lr,
                      uint64_t seed, double* losses) {
    const ApiRange range_("rbmhip_made_train");
    return guard([&] {
        RBM_REQUIRE(made && spins && n > 0 && steps >= 0 && batch > 0, "bad arguments");
        RBM_REQUIRE(batch <= (1 << 20), "batch too large");
        Ctx& c = ctx();
        made_check_device(c, made);
        if (steps == 0) return;
        const int B = batch, Bp = made_pad(B);
        made_check_extents(made_plan(made->s.N, made->s.H, n, B));
        const bool full = (int64_t)batch == n;   // the whole sample set, in order, every step
        uint64_t* ds = reinterpret_cast<uint64_t*>(c.buf("made_train_in", (size_t)n * 2 * sizeof(uint64_t)));
        c.h2d(ds, spins, (size_t)n * 2 * sizeof(uint64_t));
        float* X = reinterpret_cast<float*>(c.buf("made_x", (size_t)Bp * made->s.Np * sizeof(float)));
        int64_t* idx = full ? nullptr : reinterpret_cast<int64_t*>(c.buf("made_idx", (size_t)B * sizeof(int64_t)));
        const int tiles = made_loss_tiles(made, Bp);
        double* part = reinterpret_cast<double*>(c.buf("made_train_part", (size_t)steps * tiles * sizeof(double)));
        for (int st = 0; st < steps; st++) {
            if (!full) {
                c.launch("made_batch", [&] {
                    return launch_made_batch_index(B, n, (uint32_t)(made->t + st), (uint32_t)seed, (uint32_t)(seed >> 32), idx, c.stream);
                });
            }
            c.launch("made_gather", [&] { return launch_made_gather(Bp, B, made->s.N, made->s.Np, ds, idx, X, c.stream); });
            made_forward_backward(c, made, B, Bp, X, part + (size_t)st * tiles);
            const int t = made->t + st + 1;
            c.launch("made_adam", [&] {
                return launch_adam(made->padded, made->P(), made->G(), reinterpret_cast<float*>(made->m.p),
                                   reinterpret_cast<float*>(made->v.p), t, (float)lr, 0.9f, 0.999f, 1e-8f, c.stream);
            });
        }
        made->t += steps;
        made->inf_valid = false;
        std::vector<double> hp((size_t)steps * tiles);
        c.d2h(hp.data(), part, hp.size() * sizeof(double));
        c.sync();
        if (losses)
            for (int st = 0; st < steps; st++) {
                double sum = 0.0;
                for (int k = 0; k < tiles; k++) sum += hp[(size_t)st * tiles + k];
                losses[st] = -sum / (double)B;
            }
    });
}

int rbmhip_made_log_q(rbmhip_made* made, const uint64_t* spins, int64_t n, double* out) {
    const ApiRange range_("rbmhip_made_log_q");
    return guard([&] {
        RBM_REQUIRE(made && n >= 0 && (n == 0 || (spins && out)), "bad arguments");
        if (n == 0) return;
        Ctx& c = ctx();
        const MadeInf mi = made_inf(c, made);
        uint64_t* ds = reinterpret_cast<uint64_t*>(c.buf("in_spins", (size_t)n * 2 * sizeof(uint64_t)));
        double* lq = reinterpret_cast<double*>(c.buf("made_lq", (size_t)n * sizeof(double)));
        c.h2d(ds, spins, (size_t)n * 2 * sizeof(uint64_t));
        c.launch("made_log_q", [&] { return launch_made_log_q(mi, n, ds, lq, c.stream); });
        c.d2h(out, lq, (size_t)n * sizeof(double));
        c.sync();
    });
}

int rbmhip_made_sample(rbmhip_made* made, int64_t n, uint64_t seed, uint64_t* spins, double* log_q) {
    const ApiRange range_("rbmhip_made_sample");
    return guard([&] {
        RBM_REQUIRE(made && n >= 0 && (n == 0 || spins), "bad arguments");
        if (n == 0) return;
        Ctx& c = ctx();
        const MadeInf mi = made_inf(c, made);
        uint64_t* ds = reinterpret_cast<uint64_t*>(c.buf("made_out_spins", (size_t)n * 2 * sizeof(uint64_t)));
        double* lq = reinterpret_cast<double*>(c.buf("made_lq", (size_t)n * sizeof(double)));
        c.launch("made_sample", [&] { return launch_made_sample(mi, n, (uint32_t)seed, (uint32_t)(seed >> 32), 0u, ds, lq, c.stream); });
        c.d2h(spins, ds, (size_t)n * 2 * sizeof(uint64_t));
        if (log_q) c.d2h(log_q, lq, (size_t)n * sizeof(double));
        c.sync();
    });
}

int rbmhip_learned_last_proposals(const rbmhip_ens* ens, int64_t* n, uint64_t* spins, double* log_q, double* log_psi) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 2, "not a learned-proposal ensemble");
        if (n) *n = ens->prop_n;
        if (ens->prop_n == 0 || (!spins && !log_q && !log_psi)) return;
        Ctx& c = ctx();
        RBM_REQUIRE(c.device == ens->rec_device, "proposals were drawn on another device");
        if (spins) c.d2h(spins, ens->prop_spins.p, (size_t)ens->prop_n * 2 * sizeof(uint64_t));
        if (log_q) c.d2h(log_q, ens->prop_lq.p, (size_t)ens->prop_n * sizeof(double));
        if (log_psi) c.d2h(log_psi, ens->prop_lp.p, (size_t)ens->prop_n * sizeof(cx<double>));
        c.sync();
    });
}

int rbmhip_learned_num_runs(const rbmhip_ens* ens, int64_t* out) {
    return guard([&] {
        check_ens(ens);
        RBM_REQUIRE(ens->kind == 2 && out, "bad arguments");
        *out = ens->run_n;
    });
}

int rbmhip_made_destroy(rbmhip_made* made) {
    return guard([&] { delete made; });
}

int rbmhip_learned_create(rbmhip_made* made, uint32_t num_samples, uint32_t num_thermalization_steps,
                          uint32_t num_markov_chains, uint64_t seed, rbmhip_ens** out) {
    return guard([&] {
        RBM_REQUIRE(made && out, "bad arguments");
        RBM_REQUIRE(num_markov_chains > 0, "num_markov_chains must be > 0");
        RBM_REQUIRE(num_samples % num_markov_chains == 0, "num_samples must be a multiple of num_markov_chains");
        auto e = std::make_unique<rbmhip_ens>();
        e->kind = 2;
        e->made = made;
        e->N = made->s.N;
        e->num_samples = num_samples; e->num_sweeps = 1; e->num_therm = num_thermalization_steps;
        e->num_chains = num_markov_chains; e->seed = seed;
        *out = e.release();
    });
}

}  // extern "C"
