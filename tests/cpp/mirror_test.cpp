// C++ drop-in check: a host sum with PauliSum<N,double>'s public fields goes
// through include/paulib_hip.hpp (the GPU engine) and through the oracle's C
// ABI; the results must match bit for bit.  Built by paulib_amd/build.py,
// run by tests/test_gpu_cpp_mirror.py on the GPU box.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../include/paulib_hip.hpp"

extern "C" {
void* orc_sum_new(int as, int64_t nq);
void orc_sum_free(void*);
int64_t orc_sum_len(const void*);
int orc_sum_set(void*, int64_t, const uint64_t*, const uint64_t*, const uint8_t*, const double*);
void orc_sum_get(const void*, uint64_t*, uint64_t*, uint8_t*, double*);
int orc_multiply(void*, const void*);
int orc_canonicalize(void*);
int orc_apply_H(void*, int64_t);
int orc_apply_CNOT(void*, int64_t, int64_t);
int orc_evolve(void*, const void*);
int orc_trim(void*, int64_t);
int orc_apply_gate(void*, int, int64_t, int64_t);
int orc_apply_rot_axis(void*, int, int64_t, double);
int orc_elementwise_multiply(void*, const void*);
int orc_trim_keep(void*, int64_t, const void*);
int orc_sub(void*, const void*);
int orc_scale(void*, double);
double orc_lookup(const void*, const uint64_t*, const uint64_t*, uint8_t);
double orc_trace(const void*);
double orc_frobenius_norm2(const void*);
}

// The fields of paulib::PauliString<N> that lookup reads (include/PauliStringCore.h:24-26).
template <std::size_t N>
struct HostPauliString {
    std::array<std::size_t, (N + 63) / 64> x{}, z{};
    uint8_t flags = 0;
};

// The public fields of paulib::PauliSum<N,double> (include/PauliSumCore.h:20-45).
template <std::size_t N>
struct HostPauliSum {
    static constexpr std::size_t ArraySize = (N + 63) / 64;
    std::array<std::vector<std::size_t>, ArraySize> x_data;
    std::array<std::vector<std::size_t>, ArraySize> z_data;
    std::vector<uint8_t> flags;
    std::vector<double> value;
    std::size_t num_qubits_ = N;
    std::size_t num_qubits() const { return num_qubits_; }
    std::size_t get_length() const { return value.size(); }
};

template <std::size_t N>
HostPauliSum<N> random_sum(std::size_t n, unsigned seed) {
    std::mt19937_64 rng(seed);
    HostPauliSum<N> s;
    for (std::size_t k = 0; k < s.ArraySize; ++k) {
        s.x_data[k].resize(n);
        s.z_data[k].resize(n);
        for (std::size_t t = 0; t < n; ++t) {
            s.x_data[k][t] = rng() & rng() & rng();
            s.z_data[k][t] = rng() & rng() & rng();
        }
    }
    for (std::size_t t = 0; t < n; ++t) {
        s.flags.push_back((uint8_t)(rng() & 3));
        s.value.push_back((t % 2 ? -1.0 : 1.0) * (0.5 + (double)t));
    }
    return s;
}

template <std::size_t N>
void* to_oracle(const HostPauliSum<N>& h) {
    const int64_t n = (int64_t)h.get_length();
    constexpr int AS = (int)HostPauliSum<N>::ArraySize;
    std::vector<uint64_t> x((std::size_t)AS * n), z((std::size_t)AS * n);
    for (int k = 0; k < AS; ++k)
        for (int64_t t = 0; t < n; ++t) {
            x[(std::size_t)(k * n + t)] = h.x_data[k][(std::size_t)t];
            z[(std::size_t)(k * n + t)] = h.z_data[k][(std::size_t)t];
        }
    void* o = orc_sum_new(AS, (int64_t)h.num_qubits());
    orc_sum_set(o, n, x.data(), z.data(), h.flags.data(), h.value.data());
    return o;
}

template <std::size_t N>
bool same(const HostPauliSum<N>& g, void* o, const char* what) {
    constexpr int AS = (int)HostPauliSum<N>::ArraySize;
    const int64_t n = orc_sum_len(o);
    if ((int64_t)g.get_length() != n) {
        std::printf("FAIL %s: length %zu vs %lld\n", what, g.get_length(), (long long)n);
        return false;
    }
    std::vector<uint64_t> x((std::size_t)AS * n), z((std::size_t)AS * n);
    std::vector<uint8_t> f((std::size_t)n);
    std::vector<double> v((std::size_t)n);
    orc_sum_get(o, x.data(), z.data(), f.data(), v.data());
    for (int64_t t = 0; t < n; ++t) {
        for (int k = 0; k < AS; ++k)
            if (g.x_data[k][(std::size_t)t] != x[(std::size_t)(k * n + t)] ||
                g.z_data[k][(std::size_t)t] != z[(std::size_t)(k * n + t)]) {
                std::printf("FAIL %s: string at %lld\n", what, (long long)t);
                return false;
            }
        if (g.flags[(std::size_t)t] != f[(std::size_t)t] || std::memcmp(&g.value[(std::size_t)t], &v[(std::size_t)t], 8)) {
            std::printf("FAIL %s: phase/value at %lld\n", what, (long long)t);
            return false;
        }
    }
    std::printf("ok   %s (%lld terms)\n", what, (long long)n);
    return true;
}

template <std::size_t N>
bool run_all() {
    bool ok = true;
    auto A = random_sum<N>(300, 1), B = random_sum<N>(200, 2);
    // product + canonicalize through the device sum API
    paulib_hip::DevicePauliSum<N> dA(A), dB(B);
    auto prod = dA.multiply_canonical(dB).template download<HostPauliSum<N>>();
    void* oa = to_oracle(A);
    void* ob = to_oracle(B);
    orc_multiply(oa, ob);
    void* raw = to_oracle(A);
    orc_multiply(raw, ob);
    orc_canonicalize(oa);
    ok &= same(prod, oa, "multiply_canonical");
    // uncombined product in place on a host sum (the one-line drop-in)
    auto H = A;
    paulib_hip::multiply_on_gpu<N>(H, B);
    ok &= same(H, raw, "multiply_on_gpu");
    // canonicalize + Clifford + evolve on a device sum
    paulib_hip::DevicePauliSum<N> d(A);
    d.apply_H(3);
    d.apply_CNOT(1, N - 1);
    HostPauliSum<N> rot;
    for (std::size_t k = 0; k < rot.ArraySize; ++k) {
        rot.x_data[k] = {k == 0 ? 0x5ull : 0ull};
        rot.z_data[k] = {k == 0 ? 0x6ull : 0ull};
    }
    rot.flags = {0};
    rot.value = {0.7};
    d.evolve(rot);
    d.canonicalize();
    void* oc = to_oracle(A);
    orc_apply_H(oc, 3);
    orc_apply_CNOT(oc, 1, (int64_t)N - 1);
    void* orot = to_oracle(rot);
    orc_evolve(oc, orot);
    orc_canonicalize(oc);
    ok &= same(d.template download<HostPauliSum<N>>(), oc, "H/CNOT/evolve/canonicalize");
    // trim: the largest |c| in term order
    d.trim(77);
    orc_trim(oc, 77);
    ok &= same(d.template download<HostPauliSum<N>>(), oc, "trim");
    // the extended gate set, half-angle rotations and elementwise_multiply
    d.apply_sqrt_X(2);
    d.apply_YC_Z(0, N - 2);
    d.apply_iSWAP(N - 1, 4);
    d.apply_RY(5, 0.3);
    d.apply_RZ(N - 1, -1.1);
    orc_apply_gate(oc, 8, 2, 0);
    orc_apply_gate(oc, 17, 0, (int64_t)N - 2);
    orc_apply_gate(oc, 20, (int64_t)N - 1, 4);
    orc_apply_rot_axis(oc, 2, 5, 0.3);
    orc_apply_rot_axis(oc, 3, (int64_t)N - 1, -1.1);
    ok &= same(d.template download<HostPauliSum<N>>(), oc, "gates/RY/RZ");
    // trim(max_terms, keep), operator-=, scalar *, lookup, trace, frobenius_norm2
    {
        auto K = random_sum<N>(40, 9);
        for (std::size_t k = 0; k < K.ArraySize; ++k)
            for (std::size_t t = 0; t < 20; ++t) {
                K.x_data[k][t] = A.x_data[k][t * 7];
                K.z_data[k][t] = A.z_data[k][t * 7];
            }
        paulib_hip::DevicePauliSum<N> t(A), dk(K);
        t.trim(50, dk);
        void* ot = to_oracle(A);
        void* ok_ = to_oracle(K);
        orc_trim_keep(ot, 50, ok_);
        ok &= same(t.template download<HostPauliSum<N>>(), ot, "trim(max_terms, keep)");
        paulib_hip::DevicePauliSum<N> u(A), w(B);
        u -= w;
        auto v2 = u * 2.5;
        void* ou = to_oracle(A);
        orc_sub(ou, ob);
        orc_scale(ou, 2.5);
        ok &= same(v2.template download<HostPauliSum<N>>(), ou, "operator-= and * c");
        HostPauliString<N> q;
        for (std::size_t k = 0; k < K.ArraySize; ++k) {
            q.x[k] = A.x_data[k][5];
            q.z[k] = A.z_data[k][5];
        }
        q.flags = A.flags[5];
        void* oA = to_oracle(A);
        const double lg = paulib_hip::DevicePauliSum<N>(A).lookup(q);
        const double lw = orc_lookup(oA, reinterpret_cast<const uint64_t*>(q.x.data()),
                                     reinterpret_cast<const uint64_t*>(q.z.data()), q.flags);
        const double ng = paulib_hip::DevicePauliSum<N>(A).frobenius_norm2(), nw = orc_frobenius_norm2(oA);
        const double tg = paulib_hip::DevicePauliSum<N>(A).trace(), tw = orc_trace(oA);
        const bool qok = lg == lw && lw != 0.0 && std::fabs(ng - nw) <= 1e-12 * nw && tg == tw;
        std::printf("%s lookup/frobenius_norm2/trace (%g %g, %g %g, %g %g)\n", qok ? "ok  " : "FAIL", lg, lw, ng,
                    nw, tg, tw);
        ok &= qok;
        for (void* p : {ot, ok_, ou, oA}) orc_sum_free(p);
    }
    auto B3 = random_sum<N>(300, 3);
    paulib_hip::DevicePauliSum<N> ea(A), eb(B3);
    ea.elementwise_multiply(eb);
    void* oe = to_oracle(A);
    void* ob3 = to_oracle(B3);
    orc_elementwise_multiply(oe, ob3);
    ok &= same(ea.template download<HostPauliSum<N>>(), oe, "elementwise_multiply");
    orc_sum_free(oe);
    orc_sum_free(ob3);
    for (void* p : {oa, ob, raw, oc, orot}) orc_sum_free(p);
    return ok;
}

int main() {
    bool ok = run_all<64>() && run_all<128>() && run_all<500>();
    // errors surface as std::runtime_error like the reference's PAULIB_GPU_CK
    try {
        paulib_hip::DevicePauliSum<64> e;
        e.apply_H(64);
        std::printf("FAIL no exception for an out-of-range qubit\n");
        ok = false;
    } catch (const std::runtime_error& ex) {
        std::printf("ok   exception: %s\n", ex.what());
    }
    std::printf(ok ? "ALL OK\n" : "FAILED\n");
    return ok ? 0 : 1;
}
