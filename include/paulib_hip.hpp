// paulib_hip.hpp -- C++ host side of the MI355X PauliSum engine.
//
// Header-only; links against libpaulib_hip.so through the C ABI in
// paulib_hip.h.  It mirrors the reference's hot-path API:
//
//   * DevicePauliSum<N>  replaces GpuPauliSum<N,double>
//     (include/gpu/GpuPauliSum.h:34-148): HBM-resident terms, upload /
//     download of any host sum with PauliSum<N,T>'s public fields
//     (x_data[k], z_data[k] as std::vector<size_t>, flags, value,
//     num_qubits_; include/PauliSumCore.h:20-45), move-only, errors thrown as
//     std::runtime_error like PAULIB_GPU_CK (:25-30).
//   * the PauliSum member operations of the path as DevicePauliSum methods
//     with the reference's names (multiply, sort_and_combine, canonicalize,
//     evolve, apply_H/S/CNOT and every other gate of PauliSumClifford.h,
//     apply_RX/RY/RZ, elementwise_multiply, truncate_weight, cutoff, trim
//     (with and without keep), scale, lookup, trace, frobenius_norm2,
//     zero_state_expectation, +=, -=, *= c, binary and unary -, * c) and
//     the free functions commutator,
//     i_commutator, apply_truncation, propagate, propagate_expectation
//     (include/PauliSumOperations.h:348-356, include/PauliPropagation.h).
//   * gpu_multiply / gpu_commutator with the reference wrapper's signatures
//     (include/gpu/GpuPauliSum.h:156-238), except that the commutator's
//     output order is deterministic.
//   * in-place host helpers (multiply_on_gpu, multiply_canonical_on_gpu,
//     canonicalize_on_gpu, sort_and_combine_on_gpu) that take the caller's
//     host PauliSum, run on the device and write the result back -- the
//     one-line change a caller makes to move an existing call onto the GPU.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "paulib_hip.h"

namespace paulib_hip {

inline void check(int rc, const char* what) {
    if (rc != PLH_OK) {
        const char* m = plh_last_error();
        throw std::runtime_error(std::string("HIP engine: ") + what + ": " + (m ? m : ""));
    }
}

class Context {
   public:
    explicit Context(int device = 0, void* stream = nullptr) { check(plh_ctx_create(device, stream, &h_), "plh_ctx_create"); }
    ~Context() { plh_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    plh_ctx* get() const { return h_; }
    void synchronize() { check(plh_ctx_synchronize(h_), "plh_ctx_synchronize"); }
    static Context& default_context() {
        static Context c(0);
        return c;
    }

   private:
    plh_ctx* h_ = nullptr;
};

template <typename T>
struct TruncationOptions {  // include/PauliPropagation.h:26-31
    std::size_t max_weight = std::numeric_limits<std::size_t>::max();
    T coefficient_cutoff = T{0};
    std::size_t max_terms = std::numeric_limits<std::size_t>::max();
    int64_t weight_arg() const {
        return max_weight == std::numeric_limits<std::size_t>::max() ? -1 : (int64_t)max_weight;
    }
    int64_t terms_arg() const {
        return max_terms == std::numeric_limits<std::size_t>::max() ? -1 : (int64_t)max_terms;
    }
};

template <std::size_t N, typename T = double>
class DevicePauliSum {
    static_assert(std::is_same_v<T, double>, "the engine computes coefficients in FP64");

   public:
    static constexpr int AS = (int)((N + 63) / 64);

    explicit DevicePauliSum(Context& ctx = Context::default_context(), int64_t capacity = 16,
                            std::size_t num_qubits = N)
        : ctx_(&ctx) {
        check(plh_sum_create(ctx.get(), AS, (int64_t)num_qubits, capacity, &h_), "plh_sum_create");
    }
    template <class HostSum>
    explicit DevicePauliSum(const HostSum& src, Context& ctx = Context::default_context())
        : DevicePauliSum(ctx, (int64_t)src.get_length(), src.num_qubits()) {
        upload(src);
    }
    ~DevicePauliSum() {
        if (h_) plh_sum_destroy(h_);
    }
    DevicePauliSum(const DevicePauliSum&) = delete;
    DevicePauliSum& operator=(const DevicePauliSum&) = delete;
    DevicePauliSum(DevicePauliSum&& o) noexcept : ctx_(o.ctx_), h_(o.h_) { o.h_ = nullptr; }
    DevicePauliSum& operator=(DevicePauliSum&& o) noexcept {
        if (this != &o) {
            if (h_) plh_sum_destroy(h_);
            ctx_ = o.ctx_;
            h_ = o.h_;
            o.h_ = nullptr;
        }
        return *this;
    }

    int64_t size() const { return h_ ? plh_sum_length(h_) : 0; }
    std::size_t get_length() const { return (std::size_t)size(); }
    std::size_t num_qubits() const { return (std::size_t)plh_sum_num_qubits(h_); }
    plh_sum* handle() const { return h_; }
    Context& context() const { return *ctx_; }

    // Pack the AS host vectors of src plane by plane (GpuPauliSum::upload).
    template <class HostSum>
    void upload(const HostSum& src) {
        const int64_t n = (int64_t)src.get_length();
        std::vector<uint64_t> px((std::size_t)AS * n), pz((std::size_t)AS * n);
        for (int k = 0; k < AS; ++k) {
            if (n) {
                std::memcpy(px.data() + (std::size_t)k * n, src.x_data[k].data(), (std::size_t)n * 8);
                std::memcpy(pz.data() + (std::size_t)k * n, src.z_data[k].data(), (std::size_t)n * 8);
            }
        }
        check(plh_sum_upload(h_, n, px.data(), pz.data(), src.flags.data(), src.value.data()), "plh_sum_upload");
        check(plh_sum_set_num_qubits(h_, (int64_t)src.num_qubits()), "plh_sum_set_num_qubits");
    }

    template <class HostSum>
    void download(HostSum& dst) const {
        const int64_t n = size();
        std::vector<uint64_t> px((std::size_t)AS * n), pz((std::size_t)AS * n);
        dst.flags.resize((std::size_t)n);
        dst.value.resize((std::size_t)n);
        check(plh_sum_download(h_, px.data(), pz.data(), dst.flags.data(), dst.value.data()), "plh_sum_download");
        for (int k = 0; k < AS; ++k) {
            dst.x_data[k].resize((std::size_t)n);
            dst.z_data[k].resize((std::size_t)n);
            if (n) {
                std::memcpy(dst.x_data[k].data(), px.data() + (std::size_t)k * n, (std::size_t)n * 8);
                std::memcpy(dst.z_data[k].data(), pz.data() + (std::size_t)k * n, (std::size_t)n * 8);
            }
        }
        dst.num_qubits_ = num_qubits();
    }
    template <class HostSum>
    HostSum download() const {
        HostSum h;
        download(h);
        return h;
    }

    // ---- PauliSum member operations of the path
    void multiply(const DevicePauliSum& rhs) { check(plh_multiply(h_, rhs.h_), "plh_multiply"); }
    DevicePauliSum multiply_canonical(const DevicePauliSum& rhs) const {
        DevicePauliSum out(*ctx_, 16, num_qubits());
        check(plh_multiply_canonical(h_, rhs.h_, out.h_), "plh_multiply_canonical");
        return out;
    }
    void sort_and_combine() { check(plh_sort_and_combine(h_), "plh_sort_and_combine"); }
    void canonicalize() { check(plh_canonicalize(h_), "plh_canonicalize"); }
    T zero_state_expectation() const {
        double r = 0;
        check(plh_zero_state_expectation(h_, &r), "plh_zero_state_expectation");
        return r;
    }
    void truncate_weight(std::size_t w) { check(plh_truncate_weight(h_, (int64_t)w), "plh_truncate_weight"); }
    void cutoff(T eps) { check(plh_cutoff(h_, eps), "plh_cutoff"); }
    // PauliSum::trim(max_terms) (include/PauliSumUtils.h:330-343); ties keep the lower index.
    void trim(std::size_t max_terms) { check(plh_trim(h_, (int64_t)max_terms), "plh_trim"); }
    // trim(max_terms, keep) (:345-375): keep's (x, z) patterns are never dropped.
    void trim(std::size_t max_terms, const DevicePauliSum& keep) {
        check(plh_trim_keep(h_, (int64_t)max_terms, keep.h_), "plh_trim_keep");
    }
    void scale(T c) { check(plh_scale(h_, c), "plh_scale"); }  // :395-398
    // lookup (:253-268), trace (:231-243), frobenius_norm2 (:403-410).
    template <class PauliString>
    T lookup(const PauliString& ps) const {
        static_assert(sizeof(ps.x) == sizeof(uint64_t) * AS, "a Pauli string of this sum's array size");
        double r = 0;
        check(plh_lookup(h_, AS, reinterpret_cast<const uint64_t*>(ps.x.data()),
                         reinterpret_cast<const uint64_t*>(ps.z.data()), ps.flags, &r),
              "plh_lookup");
        return r;
    }
    T trace() const {
        double r = 0;
        check(plh_trace(h_, &r), "plh_trace");
        return r;
    }
    T frobenius_norm2() const {
        double r = 0;
        check(plh_frobenius_norm2(h_, &r), "plh_frobenius_norm2");
        return r;
    }
    void apply_H(std::size_t c) { check(plh_apply_H(h_, (int64_t)c), "plh_apply_H"); }
    void apply_S(std::size_t c) { check(plh_apply_S(h_, (int64_t)c), "plh_apply_S"); }
    void apply_CNOT(std::size_t c, std::size_t t) { check(plh_apply_CNOT(h_, (int64_t)c, (int64_t)t), "plh_apply_CNOT"); }
    // The rest of include/PauliSumClifford.h (:84-498), one launch each.
#define PLH_MIRROR_G1(NAME) \
    void apply_##NAME(std::size_t c) { check(plh_apply_##NAME(h_, (int64_t)c), "plh_apply_" #NAME); }
#define PLH_MIRROR_G2(NAME)                                                              \
    void apply_##NAME(std::size_t c, std::size_t t) {                                    \
        check(plh_apply_##NAME(h_, (int64_t)c, (int64_t)t), "plh_apply_" #NAME);         \
    }
    PLH_MIRROR_G1(I)
    PLH_MIRROR_G1(X)
    PLH_MIRROR_G1(Y)
    PLH_MIRROR_G1(Z)
    PLH_MIRROR_G1(H_YZ)
    PLH_MIRROR_G1(H_XY)
    PLH_MIRROR_G1(sqrt_X)
    PLH_MIRROR_G1(sqrt_Y)
    PLH_MIRROR_G2(CZ)
    PLH_MIRROR_G2(CY)
    PLH_MIRROR_G2(XC_X)
    PLH_MIRROR_G2(XC_Z)
    PLH_MIRROR_G2(XC_Y)
    PLH_MIRROR_G2(YC_X)
    PLH_MIRROR_G2(YC_Z)
    PLH_MIRROR_G2(YC_Y)
    PLH_MIRROR_G2(SWAP)
    PLH_MIRROR_G2(iSWAP)
#undef PLH_MIRROR_G1
#undef PLH_MIRROR_G2
    // apply_RX/RY/RZ (include/PauliSumOperations.h:28-104), half-angle convention.
    void apply_RX(std::size_t c, T angle) { check(plh_apply_RX(h_, (int64_t)c, angle), "plh_apply_RX"); }
    void apply_RY(std::size_t c, T angle) { check(plh_apply_RY(h_, (int64_t)c, angle), "plh_apply_RY"); }
    void apply_RZ(std::size_t c, T angle) { check(plh_apply_RZ(h_, (int64_t)c, angle), "plh_apply_RZ"); }
    // elementwise_multiply (include/PauliSumOperations.h:220-250).
    void elementwise_multiply(const DevicePauliSum& rhs) {
        check(plh_elementwise_multiply(h_, rhs.h_), "plh_elementwise_multiply");
    }
    template <class HostSum>
    void evolve(const HostSum& rotations) {
        const int64_t n = (int64_t)rotations.get_length();
        std::vector<uint64_t> px((std::size_t)AS * n), pz((std::size_t)AS * n);
        for (int64_t r = 0; r < n; ++r)
            for (int k = 0; k < AS; ++k) {
                px[(std::size_t)(r * AS + k)] = rotations.x_data[k][(std::size_t)r];
                pz[(std::size_t)(r * AS + k)] = rotations.z_data[k][(std::size_t)r];
            }
        check(plh_evolve(h_, n, px.data(), pz.data(), rotations.flags.data(), rotations.value.data()), "plh_evolve");
    }
    DevicePauliSum& operator+=(const DevicePauliSum& rhs) {
        check(plh_add(h_, rhs.h_), "plh_add");
        return *this;
    }
    // operator-= and scalar operator*= (include/PauliSumCore.h:191-250).
    DevicePauliSum& operator-=(const DevicePauliSum& rhs) {
        check(plh_sub(h_, rhs.h_), "plh_sub");
        return *this;
    }
    DevicePauliSum& operator*=(T c) {
        scale(c);
        return *this;
    }
    DevicePauliSum copy() const {
        DevicePauliSum out(*ctx_, std::max<int64_t>(size(), 16), num_qubits());
        check(plh_sum_copy(h_, out.h_), "plh_sum_copy");
        return out;
    }

   private:
    Context* ctx_;
    plh_sum* h_ = nullptr;
};

template <std::size_t N, typename T>
DevicePauliSum<N, T> commutator(const DevicePauliSum<N, T>& a, const DevicePauliSum<N, T>& b) {
    DevicePauliSum<N, T> out(a.context(), 16, a.num_qubits());
    check(plh_commutator(a.handle(), b.handle(), 0, out.handle()), "plh_commutator");
    return out;
}
template <std::size_t N, typename T>
DevicePauliSum<N, T> i_commutator(const DevicePauliSum<N, T>& a, const DevicePauliSum<N, T>& b) {
    DevicePauliSum<N, T> out(a.context(), 16, a.num_qubits());
    check(plh_commutator(a.handle(), b.handle(), 1, out.handle()), "plh_commutator");
    return out;
}

// gpu_multiply / gpu_commutator (include/gpu/GpuPauliSum.h:156-238).
template <std::size_t N, typename T>
DevicePauliSum<N, T> gpu_multiply(const DevicePauliSum<N, T>& lhs, const DevicePauliSum<N, T>& rhs) {
    if (lhs.size() > INT32_MAX || rhs.size() > INT32_MAX)
        throw std::runtime_error("gpu_multiply: input exceeds 2^31-1 terms");
    DevicePauliSum<N, T> out(lhs.context(), 16, std::max(lhs.num_qubits(), rhs.num_qubits()));
    check(plh_sum_copy(lhs.handle(), out.handle()), "plh_sum_copy");
    check(plh_sum_set_num_qubits(out.handle(), (int64_t)std::max(lhs.num_qubits(), rhs.num_qubits())), "nq");
    out.multiply(rhs);
    return out;
}
template <std::size_t N, typename T>
DevicePauliSum<N, T> gpu_commutator(const DevicePauliSum<N, T>& lhs, const DevicePauliSum<N, T>& rhs,
                                    bool multiply_by_i = false) {
    return multiply_by_i ? i_commutator(lhs, rhs) : commutator(lhs, rhs);
}

// operator- (binary and unary) and scalar operator* (include/PauliSumCore.h:197-262).
template <std::size_t N, typename T>
DevicePauliSum<N, T> operator-(const DevicePauliSum<N, T>& a, const DevicePauliSum<N, T>& b) {
    DevicePauliSum<N, T> out = a.copy();
    out -= b;
    return out;
}
template <std::size_t N, typename T>
DevicePauliSum<N, T> operator-(const DevicePauliSum<N, T>& a) {
    DevicePauliSum<N, T> out = a.copy();
    out.scale(T{-1});
    return out;
}
template <std::size_t N, typename T>
DevicePauliSum<N, T> operator*(const DevicePauliSum<N, T>& a, T c) {
    DevicePauliSum<N, T> out = a.copy();
    out *= c;
    return out;
}
template <std::size_t N, typename T>
DevicePauliSum<N, T> operator*(T c, const DevicePauliSum<N, T>& a) {
    return a * c;
}

template <std::size_t N, typename T>
void apply_truncation(DevicePauliSum<N, T>& O, const TruncationOptions<T>& o,
                      const DevicePauliSum<N, T>* keep = nullptr) {
    check(plh_apply_truncation_keep(O.handle(), o.weight_arg(), o.coefficient_cutoff, o.terms_arg(),
                                    keep ? keep->handle() : nullptr),
          "plh_apply_truncation");
}

template <std::size_t N, typename T, class HostSum>
void propagate(DevicePauliSum<N, T>& O, const HostSum& rotations, const TruncationOptions<T>& o = {}) {
    constexpr int AS = DevicePauliSum<N, T>::AS;
    const int64_t n = (int64_t)rotations.get_length();
    std::vector<uint64_t> px((std::size_t)AS * n), pz((std::size_t)AS * n);
    for (int64_t r = 0; r < n; ++r)
        for (int k = 0; k < AS; ++k) {
            px[(std::size_t)(r * AS + k)] = rotations.x_data[k][(std::size_t)r];
            pz[(std::size_t)(r * AS + k)] = rotations.z_data[k][(std::size_t)r];
        }
    check(plh_propagate(O.handle(), n, px.data(), pz.data(), rotations.flags.data(), rotations.value.data(),
                        o.weight_arg(), o.coefficient_cutoff, o.terms_arg()),
          "plh_propagate");
}

template <std::size_t N, typename T, class HostSum>
T propagate_expectation(const HostSum& O, const HostSum& rotations, const TruncationOptions<T>& o = {},
                        Context& ctx = Context::default_context()) {
    DevicePauliSum<N, T> d(O, ctx);
    propagate(d, rotations, o);
    return d.zero_state_expectation();
}

// ---- host-sum in place: the drop-in for an existing member call.
template <std::size_t N, class HostSum>
void multiply_on_gpu(HostSum& lhs, const HostSum& rhs, Context& ctx = Context::default_context()) {
    DevicePauliSum<N> a(lhs, ctx), b(rhs, ctx);
    const std::size_t nq = lhs.num_qubits();
    a.multiply(b);
    a.download(lhs);
    lhs.num_qubits_ = nq;
}
template <std::size_t N, class HostSum>
void multiply_canonical_on_gpu(HostSum& lhs, const HostSum& rhs, Context& ctx = Context::default_context()) {
    DevicePauliSum<N> a(lhs, ctx), b(rhs, ctx);
    a.multiply_canonical(b).download(lhs);
}
template <std::size_t N, class HostSum>
void canonicalize_on_gpu(HostSum& s, Context& ctx = Context::default_context()) {
    DevicePauliSum<N> a(s, ctx);
    a.canonicalize();
    a.download(s);
}
template <std::size_t N, class HostSum>
void sort_and_combine_on_gpu(HostSum& s, Context& ctx = Context::default_context()) {
    DevicePauliSum<N> a(s, ctx);
    a.sort_and_combine();
    a.download(s);
}

}  // namespace paulib_hip
