"""The oracle (and, where built, the compiled reference) against the known
answers that the reference's own tests hold for this path.  Each test names
the reference test it restates; expected strings use the reference's
get_str format ('%f * <string>')."""
import math

import numpy as np
import pytest

from helpers import cpu, oracle_kinds
from paulib_amd.host import HostSum

KINDS = oracle_kinds()


def S(terms, nq=None, as_=1):
    return HostSum.from_terms(terms, array_size=as_, num_qubits=nq)


def run(kind, host, *ops):
    s = cpu(kind, host)
    for name, *args in ops:
        getattr(s, name)(*args)
    return s.to_host()


def prod(kind, a, b):
    s = cpu(kind, a)
    s.multiply(cpu(kind, b))
    return s.to_host()


# tests/PauliSumTests.cpp:397-414 (PauliSumMul.PauliMul)
CAYLEY = {
    "I": {"I": "I", "X": "X", "Y": "Y", "Z": "Z"},
    "X": {"I": "X", "X": "I", "Y": "iZ", "Z": "-iY"},
    "Y": {"I": "Y", "X": "-iZ", "Y": "I", "Z": "iX"},
    "Z": {"I": "Z", "X": "iY", "Y": "-iX", "Z": "I"},
}


@pytest.mark.parametrize("kind", KINDS)
def test_cayley_table(kind):
    for r, row in CAYLEY.items():
        for c, res in row.items():
            out = prod(kind, S([(r, 1.0)]), S([(c, 1.0)]))
            assert out.get_str() == f"1.000000 * {res}\n", (r, c)


@pytest.mark.parametrize("kind", KINDS)
def test_single_qubit_products_flags(kind):
    # PauliSumTests.cpp:111-144: X*Y = iZ (flags 1), Y*X = -iZ (flags 3)
    xy = prod(kind, S([("X", 1.0)]), S([("Y", 1.0)]))
    assert xy.flags[0] == 1 and xy.x[0, 0] & 1 == 0 and xy.z[0, 0] & 1 == 1
    yx = prod(kind, S([("Y", 1.0)]), S([("X", 1.0)]))
    assert yx.flags[0] == 3
    # :66-109: II*II = 6 II, Z*Z = 2.25 I
    ii = prod(kind, S([("II", 2.0)]), S([("II", 3.0)]))
    assert ii.value[0] == 6.0 and ii.flags[0] == 0 and ii.x[0, 0] == 0 and ii.z[0, 0] == 0
    zz = prod(kind, S([("Z", 1.5)]), S([("Z", 1.5)]))
    assert zz.value[0] == 2.25 and zz.z[0, 0] == 0


@pytest.mark.parametrize("kind", KINDS)
def test_product_size_and_coefficients(kind):
    # :146-167
    out = prod(kind, S([("XX", 1.0), ("ZZ", 1.0), ("YY", 1.0)]), S([("XX", 1.0), ("ZZ", 1.0)]))
    assert len(out) == 6
    out = prod(kind, S([("XX", 2.0), ("ZZ", 3.0)]), S([("YY", 4.0)]))
    assert sorted(out.value.tolist()) == [8.0, 12.0]
    # GpuPauliSumTests.cpp:206-234
    out = prod(kind, S([("Z", -1.0)]), S([("Z", 3.0)]))
    assert out.value[0] == -3.0


@pytest.mark.parametrize("kind", KINDS)
def test_product_order_out_i_n_plus_j(kind):
    # PauliSumTests.cpp:421-446: out[i*n + j] = lhs[j] * rhs[i]
    a = S([("X", 2.0), ("Z", 3.0)])
    b = S([("Y", 5.0), ("I", 7.0)])
    out = prod(kind, a, b)
    assert out.value.tolist() == [10.0, 15.0, 14.0, 21.0]
    assert [out.term_string(t) for t in range(4)] == ["iZ", "-iX", "X", "Z"]


@pytest.mark.parametrize("kind", KINDS)
def test_elementary_empty_and_num_qubits(kind):
    # :456-472
    some = S([("XYZ", 1.0), ("ZZI", -0.5)])
    empty = HostSum(1, 3)
    assert len(prod(kind, some, empty)) == 0
    assert len(prod(kind, empty, some)) == 0
    a = S([("XIIIIIIIII", 1.0)], nq=10)
    out = prod(kind, a, S([("ZIIIIIIIII", 1.0)], nq=10))
    assert out.num_qubits == 10


@pytest.mark.parametrize("kind", KINDS)
def test_sort_and_combine_kats(kind):
    # :514-561
    out = run(kind, S([("XY", 2.0), ("ZZ", 1.0), ("XY", 3.0)]), ("sort_and_combine",))
    assert out.get_str() == "1.000000 * ZZ\n5.000000 * XY\n"
    out = run(kind, S([("X", 2.0), ("-X", 3.0)]), ("sort_and_combine",))
    assert out.get_str() == "2.000000 * X\n3.000000 * -X\n"
    out = run(kind, S([("X", 2.0), ("X", -2.0)]), ("sort_and_combine",))
    assert len(out) == 1 and out.value[0] == 0.0
    assert len(run(kind, HostSum(1, 1), ("sort_and_combine",))) == 0
    assert run(kind, S([("Y", 1.5)]), ("sort_and_combine",)).get_str() == "1.500000 * Y\n"
    s = S([("XY", 1.0), ("II", 4.0), ("XY", 1.0), ("ZZ", 2.0), ("II", -1.0)])
    once = run(kind, s, ("sort_and_combine",))
    twice = run(kind, once, ("sort_and_combine",))
    assert once.get_str() == twice.get_str()
    assert (once.value == twice.value).all()


@pytest.mark.parametrize("kind", KINDS)
def test_add_kats(kind):
    # :567-606 (operator+= is add + sort_and_combine)
    def add(a, b):
        s = cpu(kind, a)
        s.add(cpu(kind, b))
        return s.to_host()

    assert add(S([("XX", 1.0)]), S([("ZZ", 2.0)])).get_str() == "2.000000 * ZZ\n1.000000 * XX\n"
    assert add(S([("XX", 1.0), ("ZZ", 2.0)]), S([("XX", 3.0)])).get_str() == \
        "2.000000 * ZZ\n4.000000 * XX\n"
    a = S([("XY", 1.0), ("ZI", -2.0)])
    want = run(kind, a, ("sort_and_combine",))
    got = add(a, HostSum(1, 2))
    assert got.get_str() == want.get_str()
    # operator- (append_scaled(-1) + sort_and_combine), :608-645
    b = S([("XX", -3.0)])
    assert add(S([("XX", 1.0), ("ZZ", 2.0)]), b).get_str() == "2.000000 * ZZ\n-2.000000 * XX\n"


@pytest.mark.parametrize("kind", KINDS)
def test_canonicalize_kats(kind):
    # :1083-1111
    out = run(kind, S([("XI", 1.0), ("-XI", 0.25), ("ZI", 2.0)]), ("canonicalize",))
    terms = {out.term_string(t): out.value[t] for t in range(len(out))}
    assert terms == {"XI": 0.75, "ZI": 2.0}
    assert len(run(kind, S([("XI", 1.0), ("-XI", 1.0)]), ("canonicalize",))) == 0
    out = run(kind, S([("-iZ", 3.0)]), ("canonicalize",))
    assert out.term_string(0) == "iZ" and out.value[0] == -3.0


@pytest.mark.parametrize("kind", KINDS)
def test_zero_state_expectation_kats(kind):
    # :689-712, :1029-1037
    z = lambda h: cpu(kind, h).zero_state_expectation()
    assert z(S([("II", 1.5), ("ZI", 2.0), ("ZZ", 4.0), ("XI", 8.0), ("IY", 16.0)])) == 7.5
    assert z(S([("-Z", 3.0), ("Z", 1.0)])) == -2.0
    assert z(S([("iZ", 3.0), ("-iZ", 5.0), ("I", 1.0)])) == 1.0
    assert z(HostSum(1, 1)) == 0.0
    with_x = ["I"] * 128
    with_x[100] = "X"
    assert z(S([("".join(with_x), 5.0), ("Z" * 128, 2.0)], as_=2)) == 2.0


@pytest.mark.parametrize("kind", KINDS)
def test_hadamard_kats(kind):
    # :289-336, :390-395
    for src, dst in [("X", "Z"), ("Z", "X"), ("Y", "-Y"), ("I", "I")]:
        assert run(kind, S([(src, 1.0)]), ("apply_H", 0)).get_str() == f"1.000000 * {dst}\n"
    s = S([("IZXY", 1.0)])
    seq = ["IZXY", "IXXY", "IXZY", "-IXZY"]
    for q in range(4):
        s = run(kind, s, ("apply_H", q))
        assert s.get_str() == f"1.000000 * {seq[q]}\n"
    out = run(kind, S([("X", 2.0), ("Z", 3.0), ("Y", -1.0), ("I", 0.5)]), ("apply_H", 0))
    assert out.get_str() == "2.000000 * Z\n3.000000 * X\n-1.000000 * -Y\n0.500000 * I\n"
    out = run(kind, S([("IIIIIIIIXIIIIIII", 1.0)]), ("apply_H", 8))
    assert out.get_str() == "1.000000 * IIIIIIIIZIIIIIII\n"


CNOT_MAP = {"II": "II", "IX": "IX", "IY": "ZY", "IZ": "ZZ", "XI": "XX", "XX": "XI", "XY": "YZ",
            "XZ": "-YY", "YI": "YX", "YX": "YI", "YY": "-XZ", "YZ": "XY", "ZI": "ZI", "ZX": "ZX",
            "ZY": "IY", "ZZ": "IZ"}


@pytest.mark.parametrize("kind", KINDS)
def test_cnot_and_s_kats(kind):
    # :826-867, :882-892
    for src, dst in CNOT_MAP.items():
        assert run(kind, S([(src, 1.0)]), ("apply_CNOT", 0, 1)).get_str() == f"1.000000 * {dst}\n", src
    assert run(kind, S([("IIIIIIIIXIIIIIII", 1.0)]), ("apply_CNOT", 8, 9)).get_str() == \
        "1.000000 * IIIIIIIIXXIIIIII\n"
    assert run(kind, S([("IIIIIIIIIZIIIIII", 1.0)]), ("apply_CNOT", 8, 9)).get_str() == \
        "1.000000 * IIIIIIIIZZIIIIII\n"
    src = ["I"] * 128
    src[0] = "X"
    dst = list(src)
    dst[70] = "X"
    out = run(kind, S([("".join(src), 1.0)], as_=2), ("apply_CNOT", 0, 70))
    assert out.get_str() == "1.000000 * " + "".join(dst) + "\n"
    for a, b in {"I": "I", "X": "Y", "Y": "-X", "Z": "Z"}.items():
        assert run(kind, S([(a, 1.0)]), ("apply_S", 0)).get_str() == f"1.000000 * {b}\n"


def rot(p, th):
    return S([(p, th)])


@pytest.mark.parametrize("kind", KINDS)
def test_evolve_kats(kind):
    # :899-983
    ev = lambda th: run(kind, S([("X", 1.0)]), ("evolve", rot("Z", th)))
    assert ev(0.0).get_str() == "1.000000 * X\n"
    assert ev(math.pi).get_str() == "1.000000 * -X\n"
    assert ev(math.pi / 2).get_str() == "1.000000 * -Y\n"
    assert ev(-math.pi / 2).get_str() == "1.000000 * Y\n"
    th = math.pi / 3
    out = ev(th)
    assert len(out) == 2
    assert out.term_string(0) == "X" and out.value[0] == math.cos(th)
    assert out.term_string(1) == "Y" and out.value[1] == -math.sin(th)
    s = S([("Z", 2.0)])
    for t in [0.0, 0.3, math.pi / 2, -math.pi / 2, math.pi]:
        s = run(kind, s, ("evolve", rot("Z", t)))
    assert s.get_str() == "2.000000 * Z\n"


@pytest.mark.parametrize("kind", KINDS)
def test_evolve_fast_paths_are_limits(kind):
    # :944-975
    def signed(h, op):
        acc = 0.0
        for t in range(len(h)):
            st = h.term_string(t)
            neg = "-" in st
            if st.lstrip("-i") == op:
                acc += (-1.0 if neg else 1.0) * h.value[t]
        return acc

    for th in [0.0, math.pi / 2, -math.pi / 2, math.pi]:
        fast = run(kind, S([("X", 1.0)]), ("evolve", rot("Z", th)))
        gen = run(kind, S([("X", 1.0)]), ("evolve", rot("Z", th + 1e-7)))
        for op in ["X", "Y"]:
            assert abs(signed(fast, op) - signed(gen, op)) <= 1e-6


@pytest.mark.parametrize("kind", KINDS)
def test_filter_and_trim_kats(kind):
    # :1075-1081 term weight via truncate_weight, :1139-1147, :1113-1123
    out = run(kind, S([("IIII", 1.0), ("IXYZ", 1.0), ("ZIIZ", 1.0)]), ("truncate_weight", 2))
    assert [out.term_string(t) for t in range(len(out))] == ["IIII", "ZIIZ"]
    s = run(kind, S([("XII", 0.5), ("XYZ", 0.01), ("ZZI", -0.2)]), ("cutoff", 0.05))
    assert len(s) == 2
    s = run(kind, s, ("truncate_weight", 1))
    assert len(s) == 1 and s.term_string(0) == "XII"
    out = run(kind, S([("XI", 0.1), ("ZI", -5.0), ("IZ", 0.3), ("YI", 2.0)]), ("trim", 2))
    assert {out.term_string(t): out.value[t] for t in range(len(out))} == {"ZI": -5.0, "YI": 2.0}


def _lookup(kind, host, q):
    xw, zw, f, _ = HostSum.parse_string(q, host.array_size)
    return cpu(kind, host).lookup(xw, zw, f)


@pytest.mark.parametrize("kind", KINDS)
def test_trim_keep_kat(kind):
    # :1125-1137 (TrimProtectsKeepPatterns): XI survives, the budget left goes to ZI
    s = cpu(kind, S([("XI", 0.001), ("ZI", -5.0), ("IZ", 0.3), ("YI", 2.0)]))
    s.trim_keep(2, cpu(kind, S([("XI", 1.0)])))
    out = s.to_host()
    assert {out.term_string(t): out.value[t] for t in range(len(out))} == {"XI": 0.001, "ZI": -5.0}
    # the keep pattern ignores phase: keep '-iXI' protects 'XI'
    s = cpu(kind, S([("XI", 0.001), ("ZI", -5.0), ("IZ", 0.3), ("YI", 2.0)]))
    s.trim_keep(1, cpu(kind, S([("-iXI", 1.0)])))
    out = s.to_host()
    assert len(out) == 1 and out.term_string(0) == "XI"
    # more protected terms than max_terms: only they survive, in term order
    s = cpu(kind, S([("XI", 0.001), ("ZI", -5.0), ("IZ", 0.3), ("YI", 2.0)]))
    s.trim_keep(1, cpu(kind, S([("IZ", 1.0), ("XI", 1.0)])))
    out = s.to_host()
    assert [out.term_string(t) for t in range(len(out))] == ["XI", "IZ"]
    # n <= max_terms: untouched
    s = cpu(kind, S([("XI", 0.001), ("ZI", -5.0)]))
    s.trim_keep(2, cpu(kind, S([("YY", 1.0)])))
    assert len(s) == 2


@pytest.mark.parametrize("kind", KINDS)
def test_subtract_negate_scale_kats(kind):
    # :608-616 DisjointTermsConcatenateNegatedSorted
    s = cpu(kind, S([("XX", 1.0)]))
    s.sub(cpu(kind, S([("ZZ", 2.0)])))
    assert s.to_host().get_str() == "-2.000000 * ZZ\n1.000000 * XX\n"
    # :618-633 OverlappingTermsCombine
    s = cpu(kind, S([("XX", 1.0), ("ZZ", 2.0)]))
    s.sub(cpu(kind, S([("XX", 3.0)])))
    assert s.to_host().get_str() == "2.000000 * ZZ\n-2.000000 * XX\n"
    # :638-645 SelfSubtractionZeroesCoefficients (zeros stay)
    a = S([("XX", 1.0), ("ZZ", 2.0)])
    s = cpu(kind, a)
    s.sub(cpu(kind, a))
    out = s.to_host()
    assert len(out) == 2 and list(out.value) == [0.0, 0.0]
    # :647-657 unary minus is scale(-1); :659-683 scalar products
    s = cpu(kind, S([("XX", 1.0), ("ZZ", -2.5)]))
    s.scale(-1.0)
    assert s.to_host().get_str() == "-1.000000 * XX\n2.500000 * ZZ\n"
    s = cpu(kind, S([("XX", 1.0), ("ZZ", -2.0)]))
    s.scale(3.0)
    assert s.to_host().get_str() == "3.000000 * XX\n-6.000000 * ZZ\n"


@pytest.mark.parametrize("kind", KINDS)
def test_lookup_kats(kind):
    # tests/PauliSumTests.cpp:768-817
    s = S([("XXII", 1.0), ("IZZI", -0.5), ("IIXY", 0.25)])
    assert [_lookup(kind, s, q) for q in ("XXII", "IZZI", "IIXY")] == [1.0, -0.5, 0.25]
    s = S([("XX", 1.0), ("ZI", 2.0)])
    assert _lookup(kind, s, "IZ") == 0.0 and _lookup(kind, s, "YY") == 0.0
    assert _lookup(kind, S([], nq=2), "XX") == 0.0
    assert _lookup(kind, S([("XZ", 1.0), ("IY", 4.0), ("XZ", 2.0)]), "XZ") == 3.0
    s = S([("X", 1.0), ("-X", 0.25), ("iY", 2.0)])
    assert [_lookup(kind, s, q) for q in ("X", "-X", "iY", "-iY")] == [0.75, -0.75, 2.0, -2.0]
    s = S([("iZ", 3.0), ("Z", 1.0)])
    assert _lookup(kind, s, "Z") == 1.0 and _lookup(kind, s, "iZ") == 3.0
    a = ["I"] * 100
    b = ["I"] * 100
    a[3], a[97], b[97] = "X", "Z", "Z"
    s = S([("".join(a), 1.5), ("".join(b), -2.0)], as_=2)
    assert _lookup(kind, s, "".join(a)) == 1.5 and _lookup(kind, s, "".join(b)) == -2.0


@pytest.mark.parametrize("kind", KINDS)
def test_trace_and_norm_kats(kind):
    # :1042-1055 (trace), :1149-1155 (frobenius_norm2 = trace_product(s, s) / 2^n
    # = 0.5^2 + 0.2^2 + 1.5^2 for the canonical sum)
    assert cpu(kind, S([("IIII", 2.5), ("XIZI", 7.0), ("-IIII", 1.0), ("iIIII", 3.0)])).trace() == 16.0 * 1.5
    assert cpu(kind, S([("XYZI", 1.0), ("ZZZZ", -3.0)])).trace() == 0.0
    s = cpu(kind, S([("XII", 0.5), ("ZZI", -0.2), ("XYZ", 1.5)]))
    s.canonicalize()
    assert abs(s.frobenius_norm2() - (0.25 + 0.04 + 2.25)) <= 1e-14
