// _pyRBMonGPU: the reference's Python API (pyRBMonGPU/main.cpp:45-614) over
// the C ABI of librbmhip.so.  Class, method and property names and argument
// order follow the reference binding; arrays are numpy arrays.
//
// Differences, all deliberate:
//   * gpu=False (the reference's CPU path) is rejected: the library computes on
//     the GPU only;
//   * keyword-only extensions: Psi(..., a=None, precision=32) for a visible
//     bias and complex64/complex128 kernels, MonteCarloLoop(..., seed=0);
//   * Psi keeps beta as given (the reference's PsiBase copies alpha into beta,
//     include/quantum_state/PsiBase.hpp:176, 190);
//   * psi_O_k_vector returns the M + N*M derivatives of (b, W) (the reference
//     copies num_params values into an O_k_length array,
//     source/network_functions/PsiOkVector.cu:12-42).
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rbmhip.h"

namespace py = pybind11;
using cd = std::complex<double>;
using carr = py::array_t<cd, py::array::c_style | py::array::forcecast>;
using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using iarr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;
using uarr = py::array_t<uint32_t, py::array::c_style | py::array::forcecast>;

// Raised for RBMHIP_ERR_UNSUPPORTED and translated to NotImplementedError by
// the module's exception translator, which runs with the GIL held: check() is
// often called with the GIL released, where no Python C-API call is allowed.
struct NotImplemented : std::runtime_error {
    using std::runtime_error::runtime_error;
};

static void check(int rc) {
    if (rc == RBMHIP_OK) return;
    const std::string msg = rbmhip_last_error();
    if (rc == RBMHIP_ERR_UNSUPPORTED) throw NotImplemented(msg);
    if (rc == RBMHIP_ERR_ARG || rc == RBMHIP_ERR_NUMERIC) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

static void require_gpu(bool gpu) {
    if (!gpu)
        throw py::value_error("rbmongpu_amd computes on the GPU only: construct with gpu=True "
                              "(the reference's CPU path, gpu=False, is not provided)");
}

static std::vector<cd> to_vec(const carr& a) { return std::vector<cd>(a.data(), a.data() + a.size()); }

static carr from_vec(const std::vector<cd>& v) {
    carr out((py::ssize_t)v.size());
    std::copy(v.begin(), v.end(), out.mutable_data());
    return out;
}

// ---------------------------------------------------------------------------
struct PsiBase {
    rbmhip_psi* h = nullptr;
    int N = 0;
    bool gpu = true;
    bool free_quantum_axis = false;
    int precision = RBMHIP_FP32;
    double prefactor_ = 1.0;
    std::vector<double> alpha, beta;
    virtual ~PsiBase() { if (h) rbmhip_psi_destroy(h); }
    int64_t num_params() const { int64_t p = 0; check(rbmhip_psi_num_params(h, &p)); return p; }
    double get_prefactor() const { return prefactor_; }
    void set_prefactor(double v) { prefactor_ = v; check(rbmhip_psi_set_prefactor(h, v)); }
};

struct Psi : PsiBase {
    int M = 0;
    std::vector<cd> b, W, a;
    bool has_a = false;

    void create() {
        rbmhip_psi* p = nullptr;
        check(rbmhip_rbm_create(N, M, reinterpret_cast<const double*>(b.data()), reinterpret_cast<const double*>(W.data()),
                                has_a ? reinterpret_cast<const double*>(a.data()) : nullptr, prefactor_, precision, &p));
        if (h) rbmhip_psi_destroy(h);
        h = p;
    }
    void push() {
        check(rbmhip_rbm_set_params(h, reinterpret_cast<const double*>(b.data()), reinterpret_cast<const double*>(W.data()),
                                    has_a ? reinterpret_cast<const double*>(a.data()) : nullptr));
    }
    std::unique_ptr<Psi> copy() const {
        auto c = std::make_unique<Psi>();
        c->N = N; c->M = M; c->gpu = gpu; c->free_quantum_axis = free_quantum_axis; c->precision = precision;
        c->prefactor_ = prefactor_; c->alpha = alpha; c->beta = beta; c->b = b; c->W = W; c->a = a; c->has_a = has_a;
        c->create();
        return c;
    }
    // [alpha, beta, b, W (, a)]: source/quantum_state/Psi.cu:71-92
    carr get_params() const {
        std::vector<cd> p;
        p.reserve((size_t)num_params());
        for (double x : alpha) p.emplace_back(x, 0.0);
        for (double x : beta) p.emplace_back(x, 0.0);
        p.insert(p.end(), b.begin(), b.end());
        p.insert(p.end(), W.begin(), W.end());
        if (has_a) p.insert(p.end(), a.begin(), a.end());
        return from_vec(p);
    }
    void set_params(const carr& np) {
        if ((int64_t)np.size() != num_params()) throw py::value_error("params has the wrong length");
        const cd* p = np.data();
        for (int i = 0; i < N; i++) { alpha[i] = p[i].real(); beta[i] = p[N + i].real(); }
        std::copy(p + 2 * N, p + 2 * N + M, b.begin());
        std::copy(p + 2 * N + M, p + 2 * N + M + (size_t)N * M, W.begin());
        if (has_a) std::copy(p + 2 * N + M + (size_t)N * M, p + 2 * N + M + (size_t)N * M + N, a.begin());
        push();
    }
};

struct PsiDeep : PsiBase {
    std::vector<int32_t> sizes, conn_counts;
    std::vector<uint32_t> conns;
    std::vector<cd> weights, biases, final_weights;
    bool translational_invariance = false;
    bool ref_compat_final = true;
    unsigned N_i = 0, N_j = 0;

    void create() {
        rbmhip_psi* p = nullptr;
        check(rbmhip_deep_create(N, (int)sizes.size(), sizes.data(), conn_counts.data(), conns.data(),
                                 reinterpret_cast<const double*>(weights.data()),
                                 reinterpret_cast<const double*>(biases.data()),
                                 reinterpret_cast<const double*>(final_weights.data()), prefactor_,
                                 translational_invariance, ref_compat_final, precision, &p));
        if (h) rbmhip_psi_destroy(h);
        h = p;
    }
    std::unique_ptr<PsiDeep> copy() const {
        auto c = std::make_unique<PsiDeep>();
        c->N = N; c->gpu = gpu; c->free_quantum_axis = free_quantum_axis; c->precision = precision;
        c->prefactor_ = prefactor_; c->alpha = alpha; c->beta = beta;
        c->sizes = sizes; c->conn_counts = conn_counts; c->conns = conns; c->weights = weights; c->biases = biases;
        c->final_weights = final_weights; c->translational_invariance = translational_invariance;
        c->ref_compat_final = ref_compat_final; c->N_i = N_i; c->N_j = N_j;
        c->create();
        return c;
    }
    // [alpha, beta, (b_l, W_l) per layer, final]: source/quantum_state/PsiDeep.cu:120-140
    carr get_params() const {
        std::vector<cd> p;
        for (double x : alpha) p.emplace_back(x, 0.0);
        for (double x : beta) p.emplace_back(x, 0.0);
        size_t wo = 0, bo = 0;
        for (size_t l = 0; l < sizes.size(); l++) {
            p.insert(p.end(), biases.begin() + bo, biases.begin() + bo + sizes[l]);
            const size_t nw = (size_t)conn_counts[l] * sizes[l];
            p.insert(p.end(), weights.begin() + wo, weights.begin() + wo + nw);
            bo += sizes[l];
            wo += nw;
        }
        p.insert(p.end(), final_weights.begin(), final_weights.end());
        return from_vec(p);
    }
    void set_params(const carr& np) {
        if ((int64_t)np.size() != num_params()) throw py::value_error("params has the wrong length");
        const cd* p = np.data();
        for (int i = 0; i < N; i++) { alpha[i] = p[i].real(); beta[i] = p[N + i].real(); }
        size_t it = 2 * (size_t)N, wo = 0, bo = 0;
        for (size_t l = 0; l < sizes.size(); l++) {
            std::copy(p + it, p + it + sizes[l], biases.begin() + bo);
            it += sizes[l];
            const size_t nw = (size_t)conn_counts[l] * sizes[l];
            std::copy(p + it, p + it + nw, weights.begin() + wo);
            it += nw;
            bo += sizes[l];
            wo += nw;
        }
        std::copy(p + it, p + it + final_weights.size(), final_weights.begin());
        check(rbmhip_deep_set_params(h, reinterpret_cast<const double*>(weights.data()),
                                     reinterpret_cast<const double*>(biases.data()),
                                     reinterpret_cast<const double*>(final_weights.data())));
    }
    std::vector<carr> get_b() const {
        std::vector<carr> r;
        size_t bo = 0;
        for (size_t l = 0; l < sizes.size(); l++) {
            r.push_back(from_vec(std::vector<cd>(biases.begin() + bo, biases.begin() + bo + sizes[l])));
            bo += sizes[l];
        }
        return r;
    }
    std::vector<carr> get_W() const {
        std::vector<carr> r;
        size_t wo = 0;
        for (size_t l = 0; l < sizes.size(); l++) {
            carr a({(py::ssize_t)conn_counts[l], (py::ssize_t)sizes[l]});
            std::copy(weights.begin() + wo, weights.begin() + wo + (size_t)conn_counts[l] * sizes[l], a.mutable_data());
            r.push_back(a);
            wo += (size_t)conn_counts[l] * sizes[l];
        }
        return r;
    }
    std::vector<uarr> get_connections() const {
        std::vector<uarr> r;
        size_t wo = 0;
        for (size_t l = 0; l < sizes.size(); l++) {
            uarr a({(py::ssize_t)conn_counts[l], (py::ssize_t)sizes[l]});
            std::copy(conns.begin() + wo, conns.begin() + wo + (size_t)conn_counts[l] * sizes[l], a.mutable_data());
            r.push_back(a);
            wo += (size_t)conn_counts[l] * sizes[l];
        }
        return r;
    }
};

struct Operator {
    rbmhip_op* h = nullptr;
    bool gpu = true;
    int num_strings = 0, max_string_length = 0;
    std::vector<cd> coefficients;
    std::vector<int32_t> types, indices;
    ~Operator() { if (h) rbmhip_op_destroy(h); }
};

struct EnsBase {
    rbmhip_ens* h = nullptr;
    bool gpu = true;
    virtual ~EnsBase() { if (h) rbmhip_ens_destroy(h); }
    uint64_t num_steps() const { uint64_t n = 0; check(rbmhip_ens_num_steps(h, &n)); return n; }
};
struct MonteCarloLoop : EnsBase {};
struct ExactSummation : EnsBase {};

// Masked autoregressive proposal network (no reference counterpart).
struct MADE {
    rbmhip_made* h = nullptr;
    int N = 0, H = 0;
    ~MADE() { if (h) rbmhip_made_destroy(h); }
    int64_t num_params() const { int64_t n = 0; check(rbmhip_made_num_params(h, &n)); return n; }
};
// Independent Metropolis-Hastings with a MADE proposal; keeps its network alive.
struct LearnedMonteCarloLoop : EnsBase {
    py::object made;
};

using farr = py::array_t<float, py::array::c_style | py::array::forcecast>;
using u64arr = py::array_t<uint64_t, py::array::c_style | py::array::forcecast>;

static int64_t spin_rows(const u64arr& spins) {
    if (spins.ndim() != 2 || spins.shape(1) != 2) throw py::value_error("spins must be an (n, 2) uint64 array");
    return (int64_t)spins.shape(0);
}

// Spins (include/Spins.h:225-333), one 64-bit word as bound by the reference.
struct Spins {
    uint64_t c = 0;
    static uint64_t mask(uint64_t c, unsigned N) { return N >= 64 ? c : (c & (((uint64_t)1 << N) - 1u)); }
};

static Operator* make_operator(const carr& coefficients, const iarr& types, const iarr& indices, bool gpu) {
    require_gpu(gpu);
    if (types.ndim() != 2 || indices.ndim() != 2) throw py::value_error("pauli_types and pauli_indices must be 2-D");
    if (types.shape(0) != indices.shape(0) || types.shape(1) != indices.shape(1) || types.shape(0) != coefficients.size())
        throw py::value_error("operator table shapes do not match");
    auto op = new Operator();
    op->gpu = gpu;
    op->num_strings = (int)types.shape(0);
    op->max_string_length = (int)types.shape(1);
    op->coefficients = to_vec(coefficients);
    op->types.assign(types.data(), types.data() + types.size());
    op->indices.assign(indices.data(), indices.data() + indices.size());
    int rc = rbmhip_op_create(op->num_strings, op->max_string_length, reinterpret_cast<const double*>(op->coefficients.data()),
                              op->types.data(), op->indices.data(), &op->h);
    if (rc != RBMHIP_OK) { delete op; check(rc); }
    return op;
}

template <class P, class E>
static cd expectation(const P& psi, const Operator& op, E& ens) {
    double out[2];
    py::gil_scoped_release nogil;
    check(rbmhip_expectation(psi.h, op.h, ens.h, out));
    return cd(out[0], out[1]);
}

template <class P, class E>
static std::vector<cd> expectation_list(const P& psi, const std::vector<Operator*>& ops, E& ens) {
    std::vector<rbmhip_op*> hs;
    for (auto* o : ops) hs.push_back(o->h);
    std::vector<double> out(2 * ops.size());
    {
        py::gil_scoped_release nogil;
        check(rbmhip_expectation_list(psi.h, hs.data(), (int)hs.size(), ens.h, out.data()));
    }
    std::vector<cd> r;
    for (size_t i = 0; i < ops.size(); i++) r.emplace_back(out[2 * i], out[2 * i + 1]);
    return r;
}

static int estimator_code(const std::string& name) {
    if (name == "auto") return RBMHIP_GRAD_AUTO;
    if (name == "E_k") return RBMHIP_GRAD_EK;
    if (name == "hermitian") return RBMHIP_GRAD_HERMITIAN;
    throw py::value_error("estimator must be 'auto', 'E_k' or 'hermitian'");
}

template <class P, class E>
static std::pair<carr, cd> gradient(const P& psi, const Operator& op, E& ens, const std::string& estimator) {
    const int64_t n = psi.num_params();
    carr g((py::ssize_t)n);
    double E2[2];
    double* gp = reinterpret_cast<double*>(g.mutable_data());
    const int est = estimator_code(estimator);
    {
        py::gil_scoped_release nogil;
        check(rbmhip_gradient_estimator(psi.h, op.h, ens.h, est, gp, E2));
    }
    return {g, cd(E2[0], E2[1])};
}

template <class P, class E>
static carr fluctuation_gradient(const P& psi, const Operator& op, E& ens) {
    carr g((py::ssize_t)psi.num_params());
    double* gp = reinterpret_cast<double*>(g.mutable_data());
    {
        py::gil_scoped_release nogil;
        check(rbmhip_fluctuation_gradient(psi.h, op.h, ens.h, gp));
    }
    return g;
}

template <class P, class E>
static std::pair<cd, double> corrected(const P& psi, const Operator& op, const Operator& op2, E& ens) {
    double v[2], x = 0;
    {
        py::gil_scoped_release nogil;
        check(rbmhip_corrected(psi.h, op.h, op2.h, ens.h, v, &x));
    }
    return {cd(v[0], v[1]), x};
}

template <class P, class E>
static std::vector<cd> difference(const P& psi, const P& psi_prime, const std::vector<Operator*>& ops, E& ens) {
    std::vector<rbmhip_op*> hs;
    for (auto* o : ops) hs.push_back(o->h);
    std::vector<double> out(2 * ops.size());
    {
        py::gil_scoped_release nogil;
        check(rbmhip_difference(psi.h, psi_prime.h, hs.data(), (int)hs.size(), ens.h, out.data()));
    }
    std::vector<cd> r;
    for (size_t i = 0; i < ops.size(); i++) r.emplace_back(out[2 * i], out[2 * i + 1]);
    return r;
}

template <class P, class E>
static std::pair<double, cd> fluctuation(const P& psi, const Operator& op, E& ens) {
    double f = 0, E2[2];
    {
        py::gil_scoped_release nogil;
        check(rbmhip_fluctuation(psi.h, op.h, ens.h, &f, E2));
    }
    return {f, cd(E2[0], E2[1])};
}

struct ExpectationValue {
    bool gpu = true;
};

struct HilbertSpaceDistance {
    unsigned N = 0, num_params = 0;
    bool gpu = true;
};

// psi.free_quantum_axis selects the form with the axis angles of both
// wavefunctions (HilbertSpaceDistance.cu:199-205)
template <class P, class E>
static void hs_run(const P& psi, const P& psi_prime, const Operator& op, bool is_unitary, E& ens, int want_grad, double nu,
                   double* d, double* g) {
    py::gil_scoped_release nogil;
    if (psi.free_quantum_axis)
        check(rbmhip_hs_distance_axis(psi.h, psi_prime.h, op.h, is_unitary, ens.h, want_grad, nu, psi.alpha.data(),
                                      psi.beta.data(), psi_prime.alpha.data(), psi_prime.beta.data(), d, g));
    else
        check(rbmhip_hs_distance(psi.h, psi_prime.h, op.h, is_unitary, ens.h, want_grad, nu, d, g));
}

template <class P, class E>
static double hs_call(HilbertSpaceDistance& hs, const P& psi, const P& psi_prime, const Operator& op, bool is_unitary, E& ens) {
    double d = 0;
    hs_run(psi, psi_prime, op, is_unitary, ens, 0, 1.0, &d, nullptr);
    return d;
}

template <class P, class E>
static std::pair<carr, double> hs_gradient(HilbertSpaceDistance& hs, const P& psi, const P& psi_prime, const Operator& op,
                                           bool is_unitary, E& ens, float nu) {
    const int64_t n = psi_prime.num_params();
    if ((int64_t)hs.num_params != n) throw py::value_error("HilbertSpaceDistance was built for a different num_params");
    carr g((py::ssize_t)n);
    double d = 0;
    hs_run(psi, psi_prime, op, is_unitary, ens, 1, (double)nu, &d, reinterpret_cast<double*>(g.mutable_data()));
    return {g, d};
}

template <class P>
static carr psi_vector(const P& psi) {
    if (psi.N > 30) throw py::value_error("vector needs N <= 30");
    carr v((py::ssize_t)1 << psi.N);
    double* vp = reinterpret_cast<double*>(v.mutable_data());
    py::gil_scoped_release nogil;
    check(rbmhip_psi_vector(psi.h, vp));
    return v;
}

template <class P>
static double psi_norm(const P& psi, ExactSummation& ens) {
    double out = 0;
    py::gil_scoped_release nogil;
    check(rbmhip_psi_norm(psi.h, ens.h, &out));
    return out;
}

template <class P>
static carr o_k_vector(const P& psi, const Spins& spins) {
    const int64_t P_ = psi.num_params();
    std::vector<cd> full((size_t)P_);
    const uint64_t s[2] = {spins.c, 0};
    check(rbmhip_o_k_vector(psi.h, s, reinterpret_cast<double*>(full.data())));
    return from_vec(std::vector<cd>(full.begin() + 2 * psi.N, full.end()));
}

using warr = py::array_t<uint64_t, py::array::c_style | py::array::forcecast>;

static void check_words(const warr& w) {
    if (w.ndim() != 2 || w.shape(1) != 2) throw py::value_error("spins must be an (n, 2) uint64 array of spin words");
}

template <class P>
static carr log_psi_batch(const P& psi, const warr& spins) {
    check_words(spins);
    const int64_t n = spins.shape(0);
    carr out((py::ssize_t)n);
    double* op = reinterpret_cast<double*>(out.mutable_data());
    const uint64_t* sp = spins.data();
    py::gil_scoped_release nogil;
    check(rbmhip_log_psi(psi.h, sp, n, op));
    return out;
}

template <class P>
static carr local_energy_batch(const P& psi, const Operator& o, const warr& spins) {
    check_words(spins);
    const int64_t n = spins.shape(0);
    carr out((py::ssize_t)n);
    double* op = reinterpret_cast<double*>(out.mutable_data());
    const uint64_t* sp = spins.data();
    py::gil_scoped_release nogil;
    check(rbmhip_local_energy(psi.h, o.h, sp, n, op));
    return out;
}

PYBIND11_MODULE(_pyRBMonGPU, m) {
    m.doc() = "RBMonGPU API on MI355X (gfx950) HIP kernels";
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const NotImplemented& e) {
            PyErr_SetString(PyExc_NotImplementedError, e.what());
        }
    });

    py::class_<Psi>(m, "Psi")
        .def(py::init([](const darr& alpha, const darr& beta, const carr& b, const carr& W, double prefactor,
                         bool free_quantum_axis, bool gpu, py::object a, int precision) {
                 require_gpu(gpu);
                 if (W.ndim() != 2) throw py::value_error("W must be 2-D (N x M)");
                 auto p = std::make_unique<Psi>();
                 p->N = (int)alpha.size();
                 p->M = (int)b.size();
                 if ((int)W.shape(0) != p->N || (int)W.shape(1) != p->M) throw py::value_error("W must have shape (N, M)");
                 if ((int)beta.size() != p->N) throw py::value_error("beta must have length N");
                 p->gpu = gpu; p->free_quantum_axis = free_quantum_axis; p->precision = precision;
                 p->prefactor_ = prefactor;
                 p->alpha.assign(alpha.data(), alpha.data() + alpha.size());
                 p->beta.assign(beta.data(), beta.data() + beta.size());
                 p->b = to_vec(b);
                 p->W = to_vec(W);
                 if (!a.is_none()) {
                     carr aa = a.cast<carr>();
                     if ((int)aa.size() != p->N) throw py::value_error("a must have length N");
                     p->a = to_vec(aa);
                     p->has_a = true;
                 }
                 p->create();
                 return p;
             }),
             py::arg("alpha"), py::arg("beta"), py::arg("b"), py::arg("W"), py::arg("prefactor"),
             py::arg("free_quantum_axis"), py::arg("gpu"), py::kw_only(), py::arg("a") = py::none(),
             py::arg("precision") = 32)
        .def("copy", [](const Psi& p) { return p.copy(); })
        .def("__pos__", [](const Psi& p) { return p.copy(); })
        .def_property_readonly("vector", [](const Psi& p) { return psi_vector(p); })
        .def("norm", [](const Psi& p, ExactSummation& e) { return psi_norm(p, e); })
        .def_property("prefactor", &Psi::get_prefactor, &Psi::set_prefactor)
        .def_readonly("gpu", &Psi::gpu)
        .def_readonly("N", &Psi::N)
        .def_readonly("M", &Psi::M)
        .def_readonly("precision", &Psi::precision)
        .def_property("alpha", [](const Psi& p) { return darr((py::ssize_t)p.alpha.size(), p.alpha.data()); },
                      [](Psi& p, const darr& v) {
                          if ((int)v.size() != p.N) throw py::value_error("alpha must have length N");
                          p.alpha.assign(v.data(), v.data() + v.size());
                      })
        .def_property("beta", [](const Psi& p) { return darr((py::ssize_t)p.beta.size(), p.beta.data()); },
                      [](Psi& p, const darr& v) {
                          if ((int)v.size() != p.N) throw py::value_error("beta must have length N");
                          p.beta.assign(v.data(), v.data() + v.size());
                      })
        .def_property("b", [](const Psi& p) { return from_vec(p.b); },
                      [](Psi& p, const carr& v) {
                          if ((int)v.size() != p.M) throw py::value_error("b must have length M");
                          p.b = to_vec(v);
                          p.push();
                      })
        .def_property("W",
                      [](const Psi& p) {
                          carr a({(py::ssize_t)p.N, (py::ssize_t)p.M});
                          std::copy(p.W.begin(), p.W.end(), a.mutable_data());
                          return a;
                      },
                      [](Psi& p, const carr& v) {
                          if (v.ndim() != 2 || (int)v.shape(0) != p.N || (int)v.shape(1) != p.M)
                              throw py::value_error("W must have shape (N, M)");
                          p.W = to_vec(v);
                          p.push();
                      })
        .def_property_readonly("a", [](const Psi& p) -> py::object {
            if (!p.has_a) return py::none();
            return from_vec(p.a);
        })
        .def_property_readonly("num_params", &Psi::num_params)
        .def_property("params", &Psi::get_params, &Psi::set_params)
        .def_readonly("free_quantum_axis", &Psi::free_quantum_axis)
        .def_property_readonly("num_angles", [](const Psi& p) { return p.M; });

    py::class_<PsiDeep>(m, "PsiDeep")
        .def(py::init([](const darr& alpha, const darr& beta, const std::vector<carr>& biases,
                         const std::vector<uarr>& connections, const std::vector<carr>& W, const carr& final_weights,
                         double prefactor, bool free_quantum_axis, bool gpu, int precision, bool ref_compat_final) {
                 require_gpu(gpu);
                 if (biases.size() != W.size() || W.size() != connections.size() || W.empty())
                     throw py::value_error("biases, connections and W must be lists of equal, non-zero length");
                 auto p = std::make_unique<PsiDeep>();
                 p->N = (int)alpha.size();
                 p->gpu = gpu; p->free_quantum_axis = free_quantum_axis; p->precision = precision;
                 p->prefactor_ = prefactor; p->ref_compat_final = ref_compat_final;
                 p->alpha.assign(alpha.data(), alpha.data() + alpha.size());
                 p->beta.assign(beta.data(), beta.data() + beta.size());
                 for (size_t l = 0; l < W.size(); l++) {
                     if (W[l].ndim() != 2 || connections[l].ndim() != 2) throw py::value_error("W and connections must be 2-D");
                     const int C = (int)W[l].shape(0), size = (int)W[l].shape(1);
                     if ((int)biases[l].size() != size || (int)connections[l].shape(0) != C ||
                         (int)connections[l].shape(1) != size)
                         throw py::value_error("layer shapes do not match");
                     p->sizes.push_back(size);
                     p->conn_counts.push_back(C);
                     p->conns.insert(p->conns.end(), connections[l].data(), connections[l].data() + connections[l].size());
                     const auto w = to_vec(W[l]);
                     p->weights.insert(p->weights.end(), w.begin(), w.end());
                     const auto bb = to_vec(biases[l]);
                     p->biases.insert(p->biases.end(), bb.begin(), bb.end());
                 }
                 if ((int)final_weights.size() != p->sizes.back())
                     throw py::value_error("final_weights must match the last layer");
                 p->final_weights = to_vec(final_weights);
                 p->create();
                 return p;
             }),
             py::arg("alpha"), py::arg("beta"), py::arg("biases"), py::arg("connections"), py::arg("W"),
             py::arg("final_weights"), py::arg("prefactor"), py::arg("free_quantum_axis"), py::arg("gpu"),
             py::kw_only(), py::arg("precision") = 32, py::arg("ref_compat_final") = true)
        .def("copy", [](const PsiDeep& p) { return p.copy(); })
        .def("__pos__", [](const PsiDeep& p) { return p.copy(); })
        .def_property("prefactor", &PsiDeep::get_prefactor, &PsiDeep::set_prefactor)
        .def_property("translational_invariance", [](const PsiDeep& p) { return p.translational_invariance; },
                      [](PsiDeep& p, bool v) {
                          p.translational_invariance = v;
                          check(rbmhip_deep_set_translational_invariance(p.h, v));
                      })
        .def_readwrite("N_i", &PsiDeep::N_i)
        .def_readwrite("N_j", &PsiDeep::N_j)
        .def_readonly("gpu", &PsiDeep::gpu)
        .def_readonly("N", &PsiDeep::N)
        .def_readonly("precision", &PsiDeep::precision)
        .def_property_readonly("num_params", &PsiDeep::num_params)
        .def_property("params", &PsiDeep::get_params, &PsiDeep::set_params)
        .def_property_readonly("alpha", [](const PsiDeep& p) { return darr((py::ssize_t)p.alpha.size(), p.alpha.data()); })
        .def_property_readonly("beta", [](const PsiDeep& p) { return darr((py::ssize_t)p.beta.size(), p.beta.data()); })
        .def_property_readonly("b", &PsiDeep::get_b)
        .def_property_readonly("connections", &PsiDeep::get_connections)
        .def_property_readonly("W", &PsiDeep::get_W)
        .def_property_readonly("final_weights", [](const PsiDeep& p) { return from_vec(p.final_weights); })
        .def_property_readonly("_vector", [](const PsiDeep& p) { return psi_vector(p); })
        .def_property_readonly("vector", [](const PsiDeep& p) { return psi_vector(p); })
        .def("norm", [](const PsiDeep& p, ExactSummation& e) { return psi_norm(p, e); })
        .def_readonly("free_quantum_axis", &PsiDeep::free_quantum_axis);

    py::class_<Operator>(m, "Operator")
        .def(py::init(&make_operator), py::arg("coefficients"), py::arg("pauli_types"), py::arg("pauli_indices"),
             py::arg("gpu"))
        .def_readonly("gpu", &Operator::gpu)
        .def_readonly("num_strings", &Operator::num_strings)
        .def_readonly("max_string_length", &Operator::max_string_length)
        .def_property_readonly("coefficients", [](const Operator& o) { return from_vec(o.coefficients); })
        .def_property_readonly("pauli_types",
                               [](const Operator& o) {
                                   iarr a({(py::ssize_t)o.num_strings, (py::ssize_t)o.max_string_length});
                                   std::copy(o.types.begin(), o.types.end(), a.mutable_data());
                                   return a;
                               })
        .def_property_readonly("pauli_indices",
                               [](const Operator& o) {
                                   iarr a({(py::ssize_t)o.num_strings, (py::ssize_t)o.max_string_length});
                                   std::copy(o.indices.begin(), o.indices.end(), a.mutable_data());
                                   return a;
                               })
        .def_property_readonly("is_hermitian",
                               [](const Operator& o) {
                                   int herm = 0;
                                   check(rbmhip_op_info(o.h, nullptr, nullptr, &herm, nullptr));
                                   return herm != 0;
                               })
        .def_property_readonly("num_flip_masks", [](const Operator& o) {
            int nd = 0, nm = 0;
            check(rbmhip_op_info(o.h, &nd, &nm, nullptr, nullptr));
            return py::make_tuple(nd, nm);
        });

    py::class_<Spins>(m, "Spins")
        .def(py::init([](uint64_t config, unsigned N) {
            Spins s;
            s.c = Spins::mask(config, N);
            return s;
        }))
        .def("array",
             [](const Spins& s, unsigned N) {
                 darr a((py::ssize_t)N);
                 for (unsigned i = 0; i < N; i++) a.mutable_data()[i] = ((s.c >> i) & 1u) ? 1.0 : -1.0;
                 return a;
             })
        .def("flip", [](const Spins& s, unsigned pos) { Spins r; r.c = s.c ^ ((uint64_t)1 << pos); return r; })
        .def("rotate_left",
             [](const Spins& s, unsigned shift, unsigned N) {
                 Spins r;
                 const uint64_t v = shift == 0 ? s.c : ((s.c << shift) | (s.c >> (N - shift)));
                 r.c = Spins::mask(v, N);
                 return r;
             })
        .def("shift_2d",
             [](const Spins& s, unsigned si, unsigned sj, unsigned nrows, unsigned ncols) {
                 // Spins.h:287-332: shift rows by si and columns by sj, periodic
                 Spins r;
                 r.c = 0;
                 for (unsigned row = 0; row < nrows; row++)
                     for (unsigned col = 0; col < ncols; col++)
                         if ((s.c >> (row * ncols + col)) & 1u)
                             r.c |= (uint64_t)1 << (((row + si) % nrows) * ncols + (col + sj) % ncols);
                 return r;
             })
        .def("hamming_distance", [](const Spins& a, const Spins& b) { return (unsigned)__builtin_popcountll(a.c ^ b.c); })
        .def_property_readonly("configuration", [](const Spins& s) { return s.c; });

    py::class_<MonteCarloLoop>(m, "MonteCarloLoop")
        .def(py::init([](unsigned num_samples, unsigned num_sweeps, unsigned num_therm, unsigned num_chains, bool gpu,
                         uint64_t seed) {
                 require_gpu(gpu);
                 auto e = std::make_unique<MonteCarloLoop>();
                 e->gpu = gpu;
                 check(rbmhip_mc_create(num_samples, num_sweeps, num_therm, num_chains, seed, &e->h));
                 return e;
             }),
             py::arg("num_samples"), py::arg("num_sweeps"), py::arg("num_thermalization_sweeps"),
             py::arg("num_markov_chains"), py::arg("gpu"), py::kw_only(), py::arg("seed") = 0)
        .def(py::init([](MonteCarloLoop& other) {
            auto e = std::make_unique<MonteCarloLoop>();
            e->gpu = other.gpu;
            check(rbmhip_mc_copy(other.h, &e->h));
            return e;
        }))
        .def("set_total_z_symmetry", [](MonteCarloLoop& e, int sector) { check(rbmhip_mc_set_total_z(e.h, sector)); })
        .def("set_fast_sweep", [](MonteCarloLoop& e, unsigned n) { check(rbmhip_mc_set_fast_sweep(e.h, n)); })
        .def_property("persistent_chains",
                      [](const MonteCarloLoop& e) {
                          int on = 0;
                          check(rbmhip_mc_get_persistent(e.h, &on));
                          return on != 0;
                      },
                      [](MonteCarloLoop& e, bool on) { check(rbmhip_mc_set_persistent(e.h, on ? 1 : 0)); },
                      "opt-in: each call after the first continues the chains from where the previous call "
                      "left them, without thermalisation (RBM wavefunctions)")
        .def("set_angle_resync", [](MonteCarloLoop& e, int sweeps) { check(rbmhip_mc_set_angle_resync(e.h, sweeps)); },
             "complex64 chains recompute theta from the configuration every `sweeps` sweeps (default 16, 0: never)")
        .def("record_angles", [](MonteCarloLoop& e, bool on) { check(rbmhip_mc_record_angles(e.h, on ? 1 : 0)); },
             "keep each chain's final angles theta (diagnostics of the incremental updates)")
        .def("last_angles", [](const MonteCarloLoop& e) {
            int64_t chains = 0;
            int M = 0;
            check(rbmhip_mc_last_angles(e.h, &chains, &M, nullptr));
            carr out({(py::ssize_t)chains, (py::ssize_t)M});
            check(rbmhip_mc_last_angles(e.h, &chains, &M, reinterpret_cast<double*>(out.mutable_data())));
            return out;
        })
        .def_property_readonly("num_steps", &MonteCarloLoop::num_steps)
        .def_property_readonly("acceptance_rate", [](const MonteCarloLoop& e) {
            uint64_t a = 0, r = 0;
            check(rbmhip_mc_acceptance(e.h, &a, &r));
            return (float)a / (float)(a + r);
        })
        .def("last_samples", [](const MonteCarloLoop& e) {
            const uint64_t n = e.num_steps();
            py::array_t<uint64_t> spins({(py::ssize_t)n, (py::ssize_t)2});
            carr lp((py::ssize_t)n), el((py::ssize_t)n);
            check(rbmhip_mc_last_samples(e.h, spins.mutable_data(), reinterpret_cast<double*>(lp.mutable_data()),
                                         reinterpret_cast<double*>(el.mutable_data())));
            return py::make_tuple(spins, lp, el);
        });

    py::class_<MADE>(m, "MADE")
        .def(py::init([](int N, int hidden, uint64_t seed, py::object params) {
                 auto md = std::make_unique<MADE>();
                 md->N = N; md->H = hidden;
                 if (params.is_none()) {
                     check(rbmhip_made_create(N, hidden, nullptr, seed, &md->h));
                 } else {
                     farr p = params.cast<farr>();
                     if (p.size() != 2 * (py::ssize_t)N * hidden + hidden + N) throw py::value_error("params has the wrong size");
                     check(rbmhip_made_create(N, hidden, p.data(), seed, &md->h));
                 }
                 return md;
             }),
             py::arg("N"), py::arg("hidden") = 256, py::kw_only(), py::arg("seed") = 0, py::arg("params") = py::none())
        .def_readonly("N", &MADE::N)
        .def_readonly("hidden", &MADE::H)
        .def_property_readonly("num_params", &MADE::num_params)
        .def_property("params",
                      [](const MADE& md) {
                          farr out((py::ssize_t)md.num_params());
                          check(rbmhip_made_get_params(md.h, out.mutable_data()));
                          return out;
                      },
                      [](MADE& md, const farr& p) {
                          if (p.size() != md.num_params()) throw py::value_error("params has the wrong size");
                          check(rbmhip_made_set_params(md.h, p.data()));
                      })
        .def("loss_and_gradient", [](MADE& md, const u64arr& spins) {
            const int64_t n = spin_rows(spins);
            farr g((py::ssize_t)md.num_params());
            double loss = 0.0;
            {
                py::gil_scoped_release nogil;
                check(rbmhip_made_loss_grad(md.h, spins.data(), n, &loss, g.mutable_data()));
            }
            return py::make_tuple(loss, g);
        })
        .def("train", [](MADE& md, const u64arr& spins, int steps, int batch, double lr, uint64_t seed) {
                 const int64_t n = spin_rows(spins);
                 darr losses((py::ssize_t)steps);
                 {
                     py::gil_scoped_release nogil;
                     check(rbmhip_made_train(md.h, spins.data(), n, steps, batch, lr, seed, losses.mutable_data()));
                 }
                 return losses;
             },
             py::arg("spins"), py::arg("steps"), py::arg("batch") = 4096, py::arg("lr") = 1e-3, py::arg("seed") = 0)
        .def("reset_optimizer", [](MADE& md) { check(rbmhip_made_reset_optimizer(md.h)); })
        .def("log_q", [](MADE& md, const u64arr& spins) {
            const int64_t n = spin_rows(spins);
            darr out((py::ssize_t)n);
            {
                py::gil_scoped_release nogil;
                check(rbmhip_made_log_q(md.h, spins.data(), n, out.mutable_data()));
            }
            return out;
        })
        .def("sample", [](MADE& md, int64_t n, uint64_t seed) {
                 py::array_t<uint64_t> spins({(py::ssize_t)n, (py::ssize_t)2});
                 darr lq((py::ssize_t)n);
                 {
                     py::gil_scoped_release nogil;
                     check(rbmhip_made_sample(md.h, n, seed, spins.mutable_data(), lq.mutable_data()));
                 }
                 return py::make_tuple(spins, lq);
             },
             py::arg("n"), py::arg("seed") = 0);

    py::class_<LearnedMonteCarloLoop>(m, "LearnedMonteCarloLoop")
        .def(py::init([](py::object made, unsigned num_samples, unsigned num_therm, unsigned num_chains, bool gpu,
                         uint64_t seed) {
                 require_gpu(gpu);
                 MADE& md = made.cast<MADE&>();
                 auto e = std::make_unique<LearnedMonteCarloLoop>();
                 e->gpu = gpu;
                 e->made = made;
                 check(rbmhip_learned_create(md.h, num_samples, num_therm, num_chains, seed, &e->h));
                 return e;
             }),
             py::arg("made"), py::arg("num_samples"), py::arg("num_thermalization_steps"), py::arg("num_markov_chains"),
             py::arg("gpu"), py::kw_only(), py::arg("seed") = 0)
        .def_property_readonly("num_steps", &LearnedMonteCarloLoop::num_steps)
        .def_property_readonly("acceptance_rate", [](const LearnedMonteCarloLoop& e) {
            uint64_t a = 0, r = 0;
            check(rbmhip_mc_acceptance(e.h, &a, &r));
            return (float)a / (float)(a + r);
        })
        .def_property_readonly("num_runs", [](const LearnedMonteCarloLoop& e) {
            int64_t n = 0;
            check(rbmhip_learned_num_runs(e.h, &n));
            return n;
        })
        .def("last_proposals", [](const LearnedMonteCarloLoop& e) {
            int64_t n = 0;
            check(rbmhip_learned_last_proposals(e.h, &n, nullptr, nullptr, nullptr));
            py::array_t<uint64_t> spins({(py::ssize_t)n, (py::ssize_t)2});
            darr lq((py::ssize_t)n);
            carr lp((py::ssize_t)n);
            check(rbmhip_learned_last_proposals(e.h, &n, spins.mutable_data(), lq.mutable_data(),
                                                reinterpret_cast<double*>(lp.mutable_data())));
            return py::make_tuple(spins, lq, lp);
        })
        .def("last_samples", [](const LearnedMonteCarloLoop& e) {
            const uint64_t n = e.num_steps();
            py::array_t<uint64_t> spins({(py::ssize_t)n, (py::ssize_t)2});
            carr lp((py::ssize_t)n), el((py::ssize_t)n);
            check(rbmhip_mc_last_samples(e.h, spins.mutable_data(), reinterpret_cast<double*>(lp.mutable_data()),
                                         reinterpret_cast<double*>(el.mutable_data())));
            return py::make_tuple(spins, lp, el);
        });

    py::class_<ExactSummation>(m, "ExactSummation")
        .def(py::init([](unsigned N, bool gpu) {
            require_gpu(gpu);
            auto e = std::make_unique<ExactSummation>();
            e->gpu = gpu;
            check(rbmhip_exact_create((int)N, &e->h));
            return e;
        }))
        .def("set_total_z_symmetry", [](ExactSummation& e, int sector) { check(rbmhip_exact_set_total_z(e.h, sector)); })
        .def_property_readonly("num_steps", &ExactSummation::num_steps);

    auto ev = py::class_<ExpectationValue>(m, "ExpectationValue");
    ev.def(py::init([](bool gpu) {
        require_gpu(gpu);
        return ExpectationValue{gpu};
    }));
#define RBM_EV_BIND(P, E)                                                                                        \
    ev.def("__call__", [](ExpectationValue&, const P& psi, const Operator& op, E& ens) { return expectation(psi, op, ens); }); \
    ev.def("__call__", [](ExpectationValue&, const P& psi, const std::vector<Operator*>& ops, E& ens) {          \
        return expectation_list(psi, ops, ens);                                                                 \
    });                                                                                                          \
    ev.def("gradient", [](ExpectationValue&, const P& psi, const Operator& op, E& ens, const std::string& est) {    \
        return gradient(psi, op, ens, est);                                                                      \
    }, py::arg("psi"), py::arg("operator_"), py::arg("spin_ensemble"), py::kw_only(), py::arg("estimator") = "auto"); \
    ev.def("fluctuation", [](ExpectationValue&, const P& psi, const Operator& op, E& ens) { return fluctuation(psi, op, ens); }); \
    ev.def("fluctuation_gradient", [](ExpectationValue&, const P& psi, const Operator& op, E& ens) {              \
        return fluctuation_gradient(psi, op, ens);                                                               \
    });
    RBM_EV_BIND(Psi, MonteCarloLoop)
    RBM_EV_BIND(Psi, ExactSummation)
    RBM_EV_BIND(PsiDeep, MonteCarloLoop)
    RBM_EV_BIND(PsiDeep, ExactSummation)
    RBM_EV_BIND(Psi, LearnedMonteCarloLoop)
    RBM_EV_BIND(PsiDeep, LearnedMonteCarloLoop)
#undef RBM_EV_BIND
    // difference is bound for Psi and corrected for PsiDeep, as in pyRBMonGPU/main.cpp:302-308
#define RBM_EV_EXTRA(E)                                                                                          \
    ev.def("difference", [](ExpectationValue&, const Psi& psi, const Psi& psi_prime, const std::vector<Operator*>& ops, E& ens) { \
        return difference(psi, psi_prime, ops, ens);                                                             \
    });                                                                                                          \
    ev.def("corrected", [](ExpectationValue&, const PsiDeep& psi, const Operator& op, const Operator& op2, E& ens) { \
        return corrected(psi, op, op2, ens);                                                                     \
    });
    RBM_EV_EXTRA(MonteCarloLoop)
    RBM_EV_EXTRA(ExactSummation)
    RBM_EV_EXTRA(LearnedMonteCarloLoop)
#undef RBM_EV_EXTRA

    auto hs = py::class_<HilbertSpaceDistance>(m, "HilbertSpaceDistance");
    hs.def(py::init([](unsigned N, unsigned num_params, bool gpu) {
        require_gpu(gpu);
        return HilbertSpaceDistance{N, num_params, gpu};
    }));
#define RBM_HS_BIND(P, E)                                                                                         \
    hs.def("__call__", &hs_call<P, E>, py::arg("psi"), py::arg("psi_prime"), py::arg("operator_"),              \
           py::arg("is_unitary"), py::arg("spin_ensemble"));                                                     \
    hs.def("gradient", &hs_gradient<P, E>, py::arg("psi"), py::arg("psi_prime"), py::arg("operator_"),          \
           py::arg("is_unitary"), py::arg("spin_ensemble"), py::arg("nu"));
    RBM_HS_BIND(Psi, MonteCarloLoop)
    RBM_HS_BIND(Psi, ExactSummation)
    RBM_HS_BIND(PsiDeep, MonteCarloLoop)
    RBM_HS_BIND(PsiDeep, ExactSummation)
    RBM_HS_BIND(Psi, LearnedMonteCarloLoop)
    RBM_HS_BIND(PsiDeep, LearnedMonteCarloLoop)
#undef RBM_HS_BIND

    m.def("log_psi_batch", &log_psi_batch<Psi>, "log psi of (n, 2) uint64 spin words");
    m.def("log_psi_batch", &log_psi_batch<PsiDeep>);
    m.def("local_energy_batch", &local_energy_batch<Psi>, "E_loc of (n, 2) uint64 spin words");
    m.def("local_energy_batch", &local_energy_batch<PsiDeep>);
    m.def("psi_O_k_vector", [](const Psi& p, const Spins& s) { return o_k_vector(p, s); });
    m.def("psi_O_k_vector", [](const PsiDeep& p, const Spins& s) { return o_k_vector(p, s); });
    m.def("get_S_matrix", [](const Psi& psi, ExactSummation& ens) {
        const int64_t P = psi.num_params();
        carr S({(py::ssize_t)P, (py::ssize_t)P});
        double* sp = reinterpret_cast<double*>(S.mutable_data());
        {
            py::gil_scoped_release nogil;
            check(rbmhip_s_matrix(psi.h, ens.h, sp));
        }
        return S;
    });
    // Stochastic reconfiguration (no reference counterpart): centred S, FP64
    // Cholesky solve, and one SR update.
#define RBM_SR_BIND(E)                                                                                            \
    m.def("get_S_matrix_centred", [](const Psi& psi, E& ens) {                                                   \
        const int64_t P = psi.num_params() - 2 * (int64_t)psi.N;                                                 \
        carr S({(py::ssize_t)P, (py::ssize_t)P});                                                                \
        double* sp = reinterpret_cast<double*>(S.mutable_data());                                                \
        {                                                                                                         \
            py::gil_scoped_release nogil;                                                                         \
            check(rbmhip_s_matrix_centred(psi.h, ens.h, sp));                                                    \
        }                                                                                                         \
        return S;                                                                                                 \
    }, py::arg("psi"), py::arg("spin_ensemble"));                                                                 \
    m.def("sr_update", [](const Psi& psi, const Operator& op, E& ens, double eps) {                              \
        carr g((py::ssize_t)psi.num_params());                                                                    \
        double e[2];                                                                                              \
        {                                                                                                         \
            py::gil_scoped_release nogil;                                                                         \
            check(rbmhip_sr_update(psi.h, op.h, ens.h, eps, reinterpret_cast<double*>(g.mutable_data()), e));    \
        }                                                                                                         \
        return py::make_tuple(g, cd(e[0], e[1]));                                                                 \
    }, py::arg("psi"), py::arg("operator_"), py::arg("spin_ensemble"), py::arg("eps") = 1e-4);
    RBM_SR_BIND(ExactSummation)
    RBM_SR_BIND(MonteCarloLoop)
    RBM_SR_BIND(LearnedMonteCarloLoop)
#undef RBM_SR_BIND
    m.def("cholesky_solve", [](const carr& A, const carr& b) {
        if (A.ndim() != 2 || A.shape(0) != A.shape(1) || b.size() != A.shape(0))
            throw py::value_error("cholesky_solve needs an n x n matrix and an n-vector");
        const int n = (int)A.shape(0);
        carr x((py::ssize_t)n);
        {
            py::gil_scoped_release nogil;
            check(rbmhip_cholesky_solve(n, reinterpret_cast<const double*>(A.data()), reinterpret_cast<const double*>(b.data()),
                                        reinterpret_cast<double*>(x.mutable_data())));
        }
        return x;
    });
    m.def("activation_function", [](cd z) {
        double in[2] = {z.real(), z.imag()}, out[2];
        check(rbmhip_activation(in, 1, 0, out));
        return cd(out[0], out[1]);
    });
    m.def("activation_derivative", [](cd z) {
        double in[2] = {z.real(), z.imag()}, out[2];
        check(rbmhip_activation(in, 1, 1, out));
        return cd(out[0], out[1]);
    });
    m.def("setDevice", [](int d) { check(rbmhip_set_device(d)); });
    m.def("start_profiling", [] { check(rbmhip_start_profiling()); });
    m.def("stop_profiling", [] { check(rbmhip_stop_profiling()); });
    m.def("device_count", [] { int n = 0; check(rbmhip_device_count(&n)); return n; });
    m.def("synchronize", [] { check(rbmhip_synchronize()); });
    m.def("timing_enable", [](bool on) { check(rbmhip_timing_enable(on)); });
    m.def("timing_reset", [] { check(rbmhip_timing_reset()); });
    m.def("timing_query", [](const std::string& name) {
        double ms = 0;
        int64_t n = 0;
        check(rbmhip_timing_query(name.c_str(), &ms, &n));
        return py::make_tuple(ms, n);
    });
}
