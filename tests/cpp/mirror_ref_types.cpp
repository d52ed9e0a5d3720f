// Compile-only check that include/paulib_hip.hpp accepts the reference's own
// types: paulib::PauliSum<N,double> for upload/download and the in-place
// helpers, paulib::PauliString<N> for lookup, a rotation list PauliSum for
// evolve/propagate.  Never linked or run (tests/test_mirror_ref_types.py
// compiles it with -fsyntax-only against the reference headers).
#include "PauliPropagation.h"
#include "PauliSum.h"

#include "../../include/paulib_hip.hpp"

template <std::size_t N>
double exercise(paulib::PauliSum<N, double>& host, const paulib::PauliSum<N, double>& rot,
                const paulib::PauliString<N>& ps) {
    using Dev = paulib_hip::DevicePauliSum<N>;
    Dev d(host), k(host);
    d.multiply(k);
    d.canonicalize();
    d.sort_and_combine();
    d.evolve(rot);
    d.trim(10);
    d.trim(10, k);
    d.scale(0.5);
    d -= k;
    d += k;
    d *= 2.0;
    Dev e = -d;
    Dev f = d - e;
    Dev g = 3.0 * f * 0.5;
    paulib_hip::TruncationOptions<double> opts;
    opts.max_weight = 8;
    opts.coefficient_cutoff = 1e-4;
    paulib_hip::apply_truncation(g, opts);
    paulib_hip::apply_truncation(g, opts, &k);
    paulib_hip::propagate(g, rot, opts);
    Dev c = paulib_hip::commutator(d, k);
    Dev ic = paulib_hip::i_commutator(d, k);
    Dev gm = paulib_hip::gpu_multiply(d, k);
    double r = g.lookup(ps) + g.trace() + g.frobenius_norm2() + g.zero_state_expectation();
    g.download(host);
    paulib_hip::multiply_on_gpu<N>(host, rot);
    paulib_hip::multiply_canonical_on_gpu<N>(host, rot);
    paulib_hip::canonicalize_on_gpu<N>(host);
    paulib_hip::sort_and_combine_on_gpu<N>(host);
    r += paulib_hip::propagate_expectation<N, double>(host, rot, opts);
    auto back = c.template download<paulib::PauliSum<N, double>>();
    return r + (double)back.get_length() + (double)ic.size() + (double)gm.size();
}

template double exercise<64>(paulib::PauliSum<64, double>&, const paulib::PauliSum<64, double>&,
                             const paulib::PauliString<64>&);
template double exercise<128>(paulib::PauliSum<128, double>&, const paulib::PauliSum<128, double>&,
                              const paulib::PauliString<128>&);
template double exercise<500>(paulib::PauliSum<500, double>&, const paulib::PauliSum<500, double>&,
                              const paulib::PauliString<500>&);
