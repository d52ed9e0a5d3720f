// ref_bridge.cpp -- C ABI over the reference's own header-only implementation.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// headers where they lie (include/PauliSum.h, include/PauliPropagation.h);
// the output goes to oracle/_ref/ and is used by the tests to pin the
// oracle's C restatement and by the benchmark as the CPU baseline.  No
// reference source is copied: this file only instantiates the templates.
//
// The ABI mirrors oracle/pauli_oracle.c (prefix ref_ instead of orc_) so the
// tests can run both through one Python wrapper.
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <omp.h>

#include "PauliSum.h"
#include "PauliPropagation.h"

namespace {

struct RefSumBase {
    int as = 1;
    virtual ~RefSumBase() = default;
    virtual int64_t len() const = 0;
    virtual int64_t nq() const = 0;
    virtual void set_nq(int64_t) = 0;
    virtual void set(int64_t n, const uint64_t*, const uint64_t*, const uint8_t*, const double*) = 0;
    virtual void get(uint64_t*, uint64_t*, uint8_t*, double*) const = 0;
    virtual RefSumBase* clone_empty() const = 0;
};

template <int AS>
struct RefSum final : RefSumBase {
    using Sum = paulib::PauliSum<64 * AS, double>;
    Sum s;
    RefSum() { as = AS; }
    int64_t len() const override { return (int64_t)s.get_length(); }
    int64_t nq() const override { return (int64_t)s.num_qubits_; }
    void set_nq(int64_t q) override { s.num_qubits_ = (std::size_t)q; }
    void set(int64_t n, const uint64_t* x, const uint64_t* z, const uint8_t* f,
             const double* v) override {
        for (int k = 0; k < AS; ++k) {
            s.x_data[k].assign(x + (int64_t)k * n, x + (int64_t)(k + 1) * n);
            s.z_data[k].assign(z + (int64_t)k * n, z + (int64_t)(k + 1) * n);
        }
        s.flags.assign(f, f + n);
        s.value.assign(v, v + n);
    }
    void get(uint64_t* x, uint64_t* z, uint8_t* f, double* v) const override {
        const int64_t n = len();
        for (int k = 0; k < AS; ++k) {
            if (n) {
                std::memcpy(x + (int64_t)k * n, s.x_data[k].data(), (size_t)n * 8);
                std::memcpy(z + (int64_t)k * n, s.z_data[k].data(), (size_t)n * 8);
            }
        }
        if (n) {
            std::memcpy(f, s.flags.data(), (size_t)n);
            std::memcpy(v, s.value.data(), (size_t)n * 8);
        }
    }
    RefSumBase* clone_empty() const override {
        auto* r = new RefSum<AS>();
        r->s.num_qubits_ = s.num_qubits_;
        return r;
    }
};

RefSumBase* make(int as) {
    switch (as) {
        case 1: return new RefSum<1>();
        case 2: return new RefSum<2>();
        case 3: return new RefSum<3>();
        case 4: return new RefSum<4>();
        case 8: return new RefSum<8>();
        case 16: return new RefSum<16>();
        default: return nullptr;
    }
}

template <class F>
int dispatch(RefSumBase* b, F&& f) {
    switch (b->as) {
        case 1: return f(static_cast<RefSum<1>*>(b)->s);
        case 2: return f(static_cast<RefSum<2>*>(b)->s);
        case 3: return f(static_cast<RefSum<3>*>(b)->s);
        case 4: return f(static_cast<RefSum<4>*>(b)->s);
        case 8: return f(static_cast<RefSum<8>*>(b)->s);
        case 16: return f(static_cast<RefSum<16>*>(b)->s);
        default: return -2;
    }
}

template <class S>
S& same(RefSumBase* b) { return static_cast<RefSum<(int)S::ArraySize>*>(b)->s; }

template <class S>
paulib::TruncationOptions<double> opts_of(int64_t max_weight, double cutoff, int64_t max_terms) {
    paulib::TruncationOptions<double> o;
    if (max_weight >= 0) o.max_weight = (std::size_t)max_weight;
    if (cutoff > 0.0) o.coefficient_cutoff = cutoff;
    if (max_terms >= 0) o.max_terms = (std::size_t)max_terms;
    return o;
}

template <class S>
void apply_gate_to(S& s, int g, std::size_t q, std::size_t r) {
    switch (g) {
        case 0: s.apply_I(q); break;
        case 1: s.apply_X(q); break;
        case 2: s.apply_Y(q); break;
        case 3: s.apply_Z(q); break;
        case 4: s.apply_H(q); break;
        case 5: s.apply_S(q); break;
        case 6: s.apply_H_YZ(q); break;
        case 7: s.apply_H_XY(q); break;
        case 8: s.apply_sqrt_X(q); break;
        case 9: s.apply_sqrt_Y(q); break;
        case 10: s.apply_CNOT(q, r); break;
        case 11: s.apply_CZ(q, r); break;
        case 12: s.apply_CY(q, r); break;
        case 13: s.apply_XC_X(q, r); break;
        case 14: s.apply_XC_Z(q, r); break;
        case 15: s.apply_XC_Y(q, r); break;
        case 16: s.apply_YC_X(q, r); break;
        case 17: s.apply_YC_Z(q, r); break;
        case 18: s.apply_YC_Y(q, r); break;
        case 19: s.apply_SWAP(q, r); break;
        default: s.apply_iSWAP(q, r); break;
    }
}

}  // namespace

extern "C" {

void* ref_sum_new(int as, int64_t nq) {
    RefSumBase* b = make(as);
    if (b) b->set_nq(nq);
    return b;
}
void ref_sum_free(void* p) { delete static_cast<RefSumBase*>(p); }
int ref_sum_as(const void* p) { return static_cast<const RefSumBase*>(p)->as; }
int64_t ref_sum_len(const void* p) { return static_cast<const RefSumBase*>(p)->len(); }
int64_t ref_sum_nq(const void* p) { return static_cast<const RefSumBase*>(p)->nq(); }
void ref_sum_set_nq(void* p, int64_t q) { static_cast<RefSumBase*>(p)->set_nq(q); }
int ref_sum_set(void* p, int64_t n, const uint64_t* x, const uint64_t* z, const uint8_t* f,
                const double* v) {
    static_cast<RefSumBase*>(p)->set(n, x, z, f, v);
    return 0;
}
void ref_sum_get(const void* p, uint64_t* x, uint64_t* z, uint8_t* f, double* v) {
    static_cast<const RefSumBase*>(p)->get(x, z, f, v);
}

int ref_multiply(void* lhs, const void* rhs) {
    auto* a = static_cast<RefSumBase*>(lhs);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(rhs));
    if (a->as != b->as) return -2;
    return dispatch(a, [&](auto& s) { s.multiply(same<std::decay_t<decltype(s)>>(b)); return 0; });
}

void* ref_commutator(const void* pa, const void* pb, int by_i) {
    auto* a = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(pa));
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(pb));
    if (a->as != b->as) return nullptr;
    RefSumBase* out = a->clone_empty();
    dispatch(a, [&](auto& s) {
        using S = std::decay_t<decltype(s)>;
        S& o = same<S>(out);
        o = by_i ? paulib::i_commutator(s, same<S>(b)) : paulib::commutator(s, same<S>(b));
        return 0;
    });
    return out;
}

int ref_sort_and_combine(void* p) {
    return dispatch(static_cast<RefSumBase*>(p), [](auto& s) { s.sort_and_combine(); return 0; });
}
int ref_canonicalize(void* p) {
    return dispatch(static_cast<RefSumBase*>(p), [](auto& s) { s.canonicalize(); return 0; });
}
int ref_add(void* lhs, const void* rhs) {
    auto* a = static_cast<RefSumBase*>(lhs);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(rhs));
    if (a->as != b->as) return -2;
    return dispatch(a, [&](auto& s) { s += same<std::decay_t<decltype(s)>>(b); return 0; });
}
double ref_zero_state_expectation(const void* p) {
    double r = 0.0;
    dispatch(const_cast<RefSumBase*>(static_cast<const RefSumBase*>(p)),
             [&](auto& s) { r = s.zero_state_expectation(); return 0; });
    return r;
}
int ref_truncate_weight(void* p, int64_t w) {
    return dispatch(static_cast<RefSumBase*>(p), [&](auto& s) { s.truncate_weight((std::size_t)w); return 0; });
}
int ref_cutoff(void* p, double eps) {
    return dispatch(static_cast<RefSumBase*>(p), [&](auto& s) { s.cutoff(eps); return 0; });
}
int ref_trim(void* p, int64_t m) {
    return dispatch(static_cast<RefSumBase*>(p), [&](auto& s) { s.trim((std::size_t)m); return 0; });
}
int ref_trim_keep(void* p, int64_t m, const void* keep) {
    auto* a = static_cast<RefSumBase*>(p);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(keep));
    if (a->as != b->as) return -2;
    return dispatch(a, [&](auto& s) { s.trim((std::size_t)m, same<std::decay_t<decltype(s)>>(b)); return 0; });
}
int ref_scale(void* p, double c) {
    return dispatch(static_cast<RefSumBase*>(p), [&](auto& s) { s.scale(c); return 0; });
}
int ref_sub(void* lhs, const void* rhs) {
    auto* a = static_cast<RefSumBase*>(lhs);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(rhs));
    if (a->as != b->as) return -2;
    return dispatch(a, [&](auto& s) { s -= same<std::decay_t<decltype(s)>>(b); return 0; });
}
double ref_lookup(const void* p, const uint64_t* x, const uint64_t* z, uint8_t flags) {
    double r = 0.0;
    dispatch(const_cast<RefSumBase*>(static_cast<const RefSumBase*>(p)), [&](auto& s) {
        using S = std::decay_t<decltype(s)>;
        paulib::PauliString<64 * S::ArraySize> ps;
        for (std::size_t k = 0; k < S::ArraySize; ++k) {
            ps.x[k] = x[k];
            ps.z[k] = z[k];
        }
        ps.flags = flags;
        r = s.lookup(ps);
        return 0;
    });
    return r;
}
double ref_trace(const void* p) {
    double r = 0.0;
    dispatch(const_cast<RefSumBase*>(static_cast<const RefSumBase*>(p)), [&](auto& s) { r = s.trace(); return 0; });
    return r;
}
double ref_frobenius_norm2(const void* p) {
    double r = 0.0;
    dispatch(const_cast<RefSumBase*>(static_cast<const RefSumBase*>(p)),
             [&](auto& s) { r = s.frobenius_norm2(); return 0; });
    return r;
}
int ref_apply_truncation_keep(void* p, int64_t max_weight, double cutoff, int64_t max_terms, const void* keep) {
    auto* a = static_cast<RefSumBase*>(p);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(keep));
    if (a->as != b->as) return -2;
    return dispatch(a, [&](auto& s) {
        paulib::apply_truncation(s, opts_of<int>(max_weight, cutoff, max_terms), &same<std::decay_t<decltype(s)>>(b));
        return 0;
    });
}
int ref_apply_H(void* p, int64_t c) {
    return dispatch(static_cast<RefSumBase*>(p), [&](auto& s) { s.apply_H((std::size_t)c); return 0; });
}
int ref_apply_S(void* p, int64_t c) {
    return dispatch(static_cast<RefSumBase*>(p), [&](auto& s) { s.apply_S((std::size_t)c); return 0; });
}
int ref_apply_CNOT(void* p, int64_t c, int64_t t) {
    return dispatch(static_cast<RefSumBase*>(p),
                    [&](auto& s) { s.apply_CNOT((std::size_t)c, (std::size_t)t); return 0; });
}
int ref_evolve(void* p, const void* rot) {
    auto* a = static_cast<RefSumBase*>(p);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(rot));
    if (a->as != b->as) return -2;
    return dispatch(a, [&](auto& s) { s.evolve(same<std::decay_t<decltype(s)>>(b)); return 0; });
}
int ref_apply_truncation(void* p, int64_t max_weight, double cutoff, int64_t max_terms) {
    return dispatch(static_cast<RefSumBase*>(p), [&](auto& s) {
        paulib::apply_truncation(s, opts_of<int>(max_weight, cutoff, max_terms));
        return 0;
    });
}
int ref_propagate(void* p, const void* rot, int64_t max_weight, double cutoff, int64_t max_terms) {
    auto* a = static_cast<RefSumBase*>(p);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(rot));
    if (a->as != b->as) return -2;
    return dispatch(a, [&](auto& s) {
        paulib::propagate(s, same<std::decay_t<decltype(s)>>(b), opts_of<int>(max_weight, cutoff, max_terms));
        return 0;
    });
}

// Mixed Clifford + rotation program, composed from the reference's methods
// the way propagate() composes evolve and apply_truncation (see
// orc_run_program in pauli_oracle.c for the op encoding).
int ref_run_program(void* p, int64_t nops, const int32_t* kind, const int64_t* q0,
                    const int64_t* q1, const uint64_t* px, const uint64_t* pz,
                    const double* theta, int64_t max_weight, double cutoff, int64_t max_terms) {
    auto* a = static_cast<RefSumBase*>(p);
    return dispatch(a, [&](auto& s) {
        using S = std::decay_t<decltype(s)>;
        constexpr int AS = (int)S::ArraySize;
        const auto opts = opts_of<int>(max_weight, cutoff, max_terms);
        for (int64_t l = nops; l-- > 0;) {
            switch (kind[l]) {
                case 0: s.apply_H((std::size_t)q0[l]); break;
                case 1: s.apply_S((std::size_t)q0[l]); break;
                case 2: s.apply_CNOT((std::size_t)q0[l], (std::size_t)q1[l]); break;
                case 3: {
                    S gate;
                    gate.num_qubits_ = s.num_qubits_;
                    paulib::PauliString<64 * AS> ps{};
                    for (int k = 0; k < AS; ++k) { ps.x[k] = px[l * AS + k]; ps.z[k] = pz[l * AS + k]; }
                    ps.num_qubits_ = (uint32_t)s.num_qubits_;
                    gate.push_back(ps, theta[l]);
                    s.evolve(gate);
                    paulib::apply_truncation(s, opts);
                    break;
                }
                default:
                    // kinds 16 + gate id: the extended gate set
                    if (kind[l] < 16 || kind[l] > 36) return -2;
                    apply_gate_to(s, kind[l] - 16, (std::size_t)q0[l], (std::size_t)q1[l]);
                    break;
            }
        }
        return 0;
    });
}

// The extended gate set by the oracle's gate ids (0 I .. 20 iSWAP); the
// argument checks are the oracle's (the reference checks none).
int ref_apply_gate(void* p, int g, int64_t c, int64_t t) {
    auto* b = static_cast<RefSumBase*>(p);
    const int64_t nqb = 64 * (int64_t)b->as;
    if (g < 0 || g > 20 || c < 0 || c >= nqb) return -2;
    if (g >= 10 && (t < 0 || t >= nqb || t == c)) return -2;
    const std::size_t q = (std::size_t)c, r = (std::size_t)t;
    return dispatch(b, [&](auto& s) { apply_gate_to(s, g, q, r); return 0; });
}

// apply_RX / apply_RY / apply_RZ by axis (1 X, 2 Y, 3 Z).
int ref_apply_rot_axis(void* p, int axis, int64_t c, double angle) {
    auto* b = static_cast<RefSumBase*>(p);
    if (axis < 1 || axis > 3 || c < 0 || c >= 64 * (int64_t)b->as) return -2;
    return dispatch(b, [&](auto& s) {
        if (axis == 1) s.apply_RX((std::size_t)c, angle);
        else if (axis == 2) s.apply_RY((std::size_t)c, angle);
        else s.apply_RZ((std::size_t)c, angle);
        return 0;
    });
}

int ref_elementwise_multiply(void* lhs, const void* rhs) {
    auto* a = static_cast<RefSumBase*>(lhs);
    auto* b = const_cast<RefSumBase*>(static_cast<const RefSumBase*>(rhs));
    if (a->as != b->as || a->len() != b->len()) return -2;
    return dispatch(a, [&](auto& s) { s.elementwise_multiply(same<std::decay_t<decltype(s)>>(b)); return 0; });
}

int ref_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
