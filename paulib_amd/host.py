"""Host-side PauliSum data: plane-major numpy SoA plus the reference's string I/O.

HostSum mirrors PauliSum<N,double>'s public fields (include/PauliSumCore.h:
20-134) in the engine's layout: x[k, t], z[k, t] (uint64, word k holds qubits
64k..64k+63, qubit q = bit q%64), flags[t] (bit0 = i, bit1 = -1), value[t].
String parsing and printing follow include/PauliStringUtils.h:40-97 and
PauliSum::get_str (include/PauliSumUtils.h:18-33).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_PREFIX = {0: "", 1: "i", 2: "-", 3: "-i"}


def words_for(num_qubits: int) -> int:
    """ArraySize = ceil(N / 64) (include/PauliStringCore.h:19-20), at least 1."""
    return max(1, (num_qubits + 63) // 64)


@dataclass
class HostSum:
    array_size: int
    num_qubits: int
    x: np.ndarray = field(default=None)  # (AS, n) uint64
    z: np.ndarray = field(default=None)
    flags: np.ndarray = field(default=None)  # (n,) uint8
    value: np.ndarray = field(default=None)  # (n,) float64

    def __post_init__(self):
        a = self.array_size
        if self.x is None:
            self.x = np.zeros((a, 0), np.uint64)
            self.z = np.zeros((a, 0), np.uint64)
            self.flags = np.zeros(0, np.uint8)
            self.value = np.zeros(0, np.float64)
        self.x = np.ascontiguousarray(self.x, dtype=np.uint64).reshape(a, -1)
        self.z = np.ascontiguousarray(self.z, dtype=np.uint64).reshape(a, -1)
        self.flags = np.ascontiguousarray(self.flags, dtype=np.uint8)
        self.value = np.ascontiguousarray(self.value, dtype=np.float64)

    def __len__(self) -> int:
        return int(self.value.shape[0])

    def copy(self) -> "HostSum":
        return HostSum(self.array_size, self.num_qubits, self.x.copy(), self.z.copy(),
                       self.flags.copy(), self.value.copy())

    # ---------------------------------------------------------- strings
    @staticmethod
    def parse_string(s: str, array_size: int):
        """PauliString(const std::string&): optional prefix -i, +i, -, +, i."""
        if not s:
            raise ValueError("Pauli string is empty")
        sign = imag = False
        pos = 0
        if s.startswith("-i"):
            sign = imag = True
            pos = 2
        elif s.startswith("+i"):
            imag = True
            pos = 2
        elif s[0] == "-":
            sign = True
            pos = 1
        elif s[0] in "+i":
            imag = s[0] == "i"
            pos = 1
        body = s[pos:]
        if len(body) > 64 * array_size:
            raise ValueError("Pauli string exceeds capacity")
        xw = [0] * array_size
        zw = [0] * array_size
        for q, ch in enumerate(body):
            w, b = divmod(q, 64)
            if ch == "X":
                xw[w] |= 1 << b
            elif ch == "Z":
                zw[w] |= 1 << b
            elif ch == "Y":
                xw[w] |= 1 << b
                zw[w] |= 1 << b
            elif ch != "I":
                raise ValueError(f"Invalid Pauli character {ch!r}")
        flags = (2 if sign else 0) | (1 if imag else 0)
        return xw, zw, flags, len(body)

    @classmethod
    def from_terms(cls, terms, array_size: int | None = None, num_qubits: int | None = None):
        """PauliSum(vector<pair<string, T>>): num_qubits = longest string."""
        terms = list(terms)
        nq = max((len(s.lstrip("+-i")) for s, _ in terms), default=0)
        if array_size is None:
            array_size = words_for(num_qubits if num_qubits is not None else nq)
        xs, zs, fs, vs = [], [], [], []
        maxq = 0
        for s, c in terms:
            xw, zw, f, q = cls.parse_string(s, array_size)
            xs.append(xw)
            zs.append(zw)
            fs.append(f)
            vs.append(float(c))
            maxq = max(maxq, q)
        n = len(terms)
        x = np.array(xs, dtype=np.uint64).reshape(n, array_size).T.copy() if n else None
        z = np.array(zs, dtype=np.uint64).reshape(n, array_size).T.copy() if n else None
        out = cls(array_size, num_qubits if num_qubits is not None else (maxq or 64 * array_size),
                  x, z, np.array(fs, np.uint8) if n else None, np.array(vs, np.float64) if n else None)
        return out

    def term_string(self, t: int) -> str:
        """PauliString::to_string over num_qubits sites with the phase prefix."""
        chars = []
        for q in range(self.num_qubits):
            w, b = divmod(q, 64)
            xb = (int(self.x[w, t]) >> b) & 1
            zb = (int(self.z[w, t]) >> b) & 1
            chars.append("Y" if xb and zb else "X" if xb else "Z" if zb else "I")
        return _PREFIX[int(self.flags[t]) & 3] + "".join(chars)

    def get_str(self) -> str:
        """PauliSum::get_str: '%f * <string>\\n' per term (std::to_string)."""
        return "".join(f"{self.value[t]:f} * {self.term_string(t)}\n" for t in range(len(self)))

    # ------------------------------------------------------- comparisons
    def key_order(self) -> np.ndarray:
        """Indices that sort the terms by (x words, z words, flags), stably."""
        keys = [self.flags]
        for k in reversed(range(self.array_size)):
            keys.append(self.z[k])
        for k in reversed(range(self.array_size)):
            keys.append(self.x[k])
        return np.lexsort(keys)
