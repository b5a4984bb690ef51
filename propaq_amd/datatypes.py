"""At-rest operator types of the drop-in API.

``PauliString`` / ``MajoranaMonomial`` and the term sums mirror the reference's
PyO3 classes (``propaq/_rust_core/__init__.pyi``; Rust in
``crates/pauli/src/string.rs:30``, ``crates/pauli/src/termsum.rs:46``,
``crates/majorana/src/monomial.rs``, ``crates/majorana/src/termsum.rs``):
same constructor arguments, properties and algebra, so observables and
generators built for propaq work unchanged.  Term sums keep their terms in a
dict keyed by the engine's interleaved key; the engine itself never sees
these objects, only packed u64 words.
"""

from __future__ import annotations

import gzip
import struct
from typing import Dict, Iterable, Iterator, Optional, Tuple

import numpy as np

_EVEN = 0x5555555555555555


def _popcount(v: int) -> int:
    return bin(v).count("1")


def _i_pow(e: int) -> complex:
    return (1, 1j, -1, -1j)[e % 4]


# ----------------------------------------------------------------- Pauli

class PauliString:
    """An n-qubit Pauli operator encoded as two integer bitmasks.

    ``x`` bit k set: qubit k has an X or Y component; ``z`` bit k set: a Z or
    Y component (00 -> I, 01 -> X, 10 -> Z, 11 -> Y).
    """

    __slots__ = ("_x", "_z", "_n")

    def __init__(self, x, z, n_qubits: int):
        x, z = int(x), int(z)
        if x < 0 or z < 0:
            raise ValueError("bitmasks must be non-negative")
        if (x | z) >> n_qubits:
            raise ValueError("bitmask has bits beyond n_qubits")
        self._x, self._z, self._n = x, z, int(n_qubits)

    @property
    def x(self) -> int:
        return self._x

    @property
    def z(self) -> int:
        return self._z

    @property
    def n_qubits(self) -> int:
        return self._n

    @property
    def weight(self) -> int:
        return _popcount(self._x | self._z)

    @property
    def key(self) -> int:
        """Interleaved engine key: bit 2q = X, bit 2q+1 = Z
        (crates/pauli/src/algebra.rs:145)."""
        k, v, q = 0, self._x | self._z, 0
        while v >> q:
            if (self._x >> q) & 1:
                k |= 1 << (2 * q)
            if (self._z >> q) & 1:
                k |= 1 << (2 * q + 1)
            q += 1
        return k

    @classmethod
    def from_key(cls, key: int, n_qubits: int) -> "PauliString":
        x = z = 0
        q = 0
        while key >> (2 * q):
            if (key >> (2 * q)) & 1:
                x |= 1 << q
            if (key >> (2 * q + 1)) & 1:
                z |= 1 << q
            q += 1
        return cls(x, z, n_qubits)

    def commutes_with(self, other: "PauliString") -> bool:
        return (_popcount(self._x & other._z) + _popcount(self._z & other._x)) % 2 == 0

    def __matmul__(self, other: "PauliString") -> Tuple[complex, "PauliString"]:
        """(phase, product) with phase in {1, i, -1, -i}."""
        a, b = self.key, other.key
        o = a ^ b
        y = lambda m: _popcount(m & (m >> 1) & int("01" * 4096, 2))
        zx = _popcount((a >> 1) & b & int("01" * 4096, 2))
        e = (y(a) + y(b) - y(o) + 2 * zx) % 4
        return _i_pow(e), PauliString.from_key(o, max(self._n, other._n))

    def trace_with_diag_state(self, diag_state) -> float:
        if self._x:
            return 0.0
        return -1.0 if _popcount(self._z & int(diag_state)) % 2 else 1.0

    def to_bytes(self) -> bytes:
        nb = max(1, (self._n + 7) // 8)
        return self._x.to_bytes(nb, "little") + self._z.to_bytes(nb, "little")

    def __hash__(self) -> int:
        return hash((self._x, self._z))

    def __eq__(self, other) -> bool:
        return isinstance(other, PauliString) and self._x == other._x and self._z == other._z

    def __repr__(self) -> str:
        chars = []
        for q in range(self._n - 1, -1, -1):
            bx, bz = (self._x >> q) & 1, (self._z >> q) & 1
            chars.append("Y" if bx and bz else ("X" if bx else ("Z" if bz else "I")))
        return f"PauliString({''.join(chars)})"


# -------------------------------------------------------------- Majorana

def _herm_exp(length: int) -> int:
    return 0 if length % 4 in (0, 1) else 1


def _resort_parity(a: int, b: int) -> int:
    count, rest = 0, b
    while rest:
        low = rest & -rest
        pos = low.bit_length() - 1
        count += _popcount(a >> (pos + 1))
        rest ^= low
    return count & 1


def jw_weight(modes: int, n_sites: int) -> int:
    """Pauli weight of the Jordan-Wigner image (majorana/src/algebra.rs:85)."""
    singles = [((modes >> (2 * s)) & 1) ^ ((modes >> (2 * s + 1)) & 1) for s in range(n_sites)]
    flip = sum(singles) & 1
    pref, w = 0, 0
    for s in range(n_sites):
        xb, yb = (modes >> (2 * s)) & 1, (modes >> (2 * s + 1)) & 1
        pref ^= singles[s]
        w += singles[s] | ((xb | yb) ^ (pref ^ flip))
    return w


class MajoranaMonomial:
    """A Majorana monomial encoded as a mode bitmask (bit k = gamma_k; site s
    holds gamma_2s and gamma_2s+1)."""

    __slots__ = ("_modes", "_n", "_np")

    def __init__(self, modes, n_modes: int, is_number_preserving: bool = True):
        modes = int(modes)
        if n_modes % 2:
            raise ValueError("n_modes must be even")
        if modes < 0 or modes >> n_modes:
            raise ValueError("mode bitmask out of range")
        self._modes, self._n, self._np = modes, int(n_modes), bool(is_number_preserving)

    @property
    def modes(self) -> int:
        return self._modes

    @property
    def n_modes(self) -> int:
        return self._n

    @property
    def is_number_preserving(self) -> bool:
        return self._np

    @property
    def length(self) -> int:
        return _popcount(self._modes)

    @property
    def weight(self) -> int:
        return jw_weight(self._modes, self._n // 2)

    @property
    def key(self) -> int:
        return self._modes

    @classmethod
    def from_key(cls, key: int, n_modes: int) -> "MajoranaMonomial":
        npres = all(((key >> (2 * s)) & 3) in (0, 3) for s in range(n_modes // 2))
        return cls(key, n_modes, npres)

    def overlap(self, other: "MajoranaMonomial") -> int:
        return _popcount(self._modes & other._modes)

    def commutes_with(self, other: "MajoranaMonomial") -> bool:
        return (self.length * other.length + self.overlap(other)) % 2 == 0

    def resulting_weight(self, other: "MajoranaMonomial") -> int:
        return jw_weight(self._modes ^ other._modes, self._n // 2)

    def __matmul__(self, other: "MajoranaMonomial") -> Tuple[complex, "MajoranaMonomial"]:
        a, b = self._modes, other._modes
        o = a ^ b
        e = (_herm_exp(_popcount(a)) + _herm_exp(_popcount(b)) - _herm_exp(_popcount(o))
             + 2 * _resort_parity(a, b)) % 4
        return _i_pow(e), MajoranaMonomial.from_key(o, max(self._n, other._n))

    def trace_with_diag_state(self, diag_state) -> float:
        d = int(diag_state)
        paired, prod = 0, 1
        for s in range(self._n // 2):
            lo, hi = (self._modes >> (2 * s)) & 1, (self._modes >> (2 * s + 1)) & 1
            if lo != hi:
                return 0.0
            if lo:
                prod *= 2 * ((d >> s) & 1) - 1
                paired += 1
        return float((1 if (paired // 2) % 2 == 0 else -1) * prod)

    def to_bytes(self) -> bytes:
        return self._modes.to_bytes(max(1, (self._n + 7) // 8), "little")

    def __hash__(self) -> int:
        return hash(self._modes)

    def __eq__(self, other) -> bool:
        return isinstance(other, MajoranaMonomial) and self._modes == other._modes

    def __repr__(self) -> str:
        return f"MajoranaMonomial(modes={self._modes:#x}, n_modes={self._n})"


# -------------------------------------------------------------- term sums

class _TermSum:
    """Dict of engine key -> coefficient with the reference's term-sum API."""

    _term_cls = None  # set by subclasses

    def __init__(self, terms: Optional[dict] = None, dtype: Optional[str] = None):
        if dtype not in (None, "float64", "float32"):
            raise ValueError('dtype must be "float64" or "float32"')
        self._dtype = dtype or "float64"
        self._terms: Dict[int, float] = {}
        self._size = 0
        if terms:
            for t, c in terms.items():
                self.add(t, c)

    # construction helpers for the engine
    @classmethod
    def _from_arrays(cls, keys: np.ndarray, coeffs: np.ndarray, size: int) -> "_TermSum":
        out = cls()
        out._size = size
        W = keys.shape[1] if keys.ndim == 2 else 1
        if W == 1:
            ks = [int(k) for k in keys.reshape(-1)]
        else:
            ks = [sum(int(w) << (64 * i) for i, w in enumerate(row)) for row in keys]
        out._terms = dict(zip(ks, (float(c) for c in coeffs)))
        return out

    def _coerce(self, c: float) -> float:
        return float(np.float32(c)) if self._dtype == "float32" else float(c)

    @property
    def dtype(self) -> str:
        return self._dtype

    @property
    def n_units(self) -> int:
        return self._size

    def add(self, term, coeff: float) -> None:
        self._size = max(self._size, self._term_size(term))
        k = term.key
        self._terms[k] = self._coerce(self._terms.get(k, 0.0) + float(coeff))

    def scale(self, factor: float) -> None:
        for k in self._terms:
            self._terms[k] = self._coerce(self._terms[k] * factor)

    def merge(self, other: "_TermSum") -> None:
        if other._dtype != self._dtype:
            raise ValueError("both term sums must share the same dtype")
        self._size = max(self._size, other._size)
        for k, c in other._terms.items():
            self._terms[k] = self._coerce(self._terms.get(k, 0.0) + c)

    def norm_squared(self) -> float:
        return float(sum(c * c for c in self._terms.values()))

    def items(self):
        return [(self._make(k), c) for k, c in self._terms.items()]

    def keys_and_coeffs(self) -> Tuple[list, list]:
        return list(self._terms.keys()), list(self._terms.values())

    def __len__(self) -> int:
        return len(self._terms)

    def __setitem__(self, term, coeff: float) -> None:
        self._size = max(self._size, self._term_size(term))
        self._terms[term.key] = self._coerce(coeff)

    def __getitem__(self, term) -> float:
        return self._terms.get(term.key, 0.0)

    def __iter__(self) -> Iterator:
        return iter(self.items())

    def copy(self) -> "_TermSum":
        out = type(self)(dtype=self._dtype)
        out._terms = dict(self._terms)
        out._size = self._size
        return out

    @property
    def sparse_key_bytes(self) -> int:
        return len(self._terms) * 8 * max(1, (2 * self._units() + 63) // 64)

    def truncate(self, policy) -> None:
        """Drop terms the policy rejects (TruncationPolicy.should_truncate)."""
        keep = {}
        for k, c in self._terms.items():
            t = self._make(k)
            if not policy.should_truncate(t.weight, abs(c)):
                keep[k] = c
        self._terms = keep

    def apply_damping(self, noise, active_modes: int) -> None:
        for k in list(self._terms):
            w = self._make(k).weight
            self._terms[k] = self._coerce(self._terms[k] * noise.damping_factor(w, active_modes))

    # gzip term file (crates/core/src/interface/term_io.rs:13-89)
    def save(self, path: str) -> None:
        save_terms(path, [(self._make(k), c) for k, c in self._terms.items()])

    @classmethod
    def from_file(cls, path: str) -> "_TermSum":
        out = cls()
        for t, c in cls._load(path):
            out.add(t, c)
        return out

    def merge_from_file(self, streamer) -> None:
        for t, c in streamer:
            self.add(t, c)


class PauliTermSum(_TermSum):
    """A weighted sum of Pauli strings with real coefficients."""

    def _term_size(self, t: PauliString) -> int:
        return t.n_qubits

    def _units(self) -> int:
        return self._size

    def _make(self, key: int) -> PauliString:
        return PauliString.from_key(key, self._size)

    @classmethod
    def _load(cls, path):
        return PauliTermStreamer.from_file(path)


class MajoranaTermSum(_TermSum):
    """A weighted sum of Majorana monomials with real coefficients."""

    def _term_size(self, t: MajoranaMonomial) -> int:
        return t.n_modes

    def _units(self) -> int:
        return self._size // 2

    def _make(self, key: int) -> MajoranaMonomial:
        return MajoranaMonomial.from_key(key, self._size)

    @classmethod
    def _load(cls, path):
        return MajoranaTermStreamer.from_file(path)


def save_terms(path: str, terms) -> None:
    """u64 n_terms, u64 key_stride, u64 system_size, then key bytes + f64 per
    term, all little-endian, gzip-compressed (term_io.rs:13-24)."""
    with gzip.open(path, "wb") as f:
        f.write(struct.pack("<Q", len(terms)))
        if terms:
            first = terms[0][0]
            stride = len(first.to_bytes())
            size = first.n_qubits if isinstance(first, PauliString) else first.n_modes
        else:
            stride, size = 0, 0
        f.write(struct.pack("<QQ", stride, size))
        for t, c in terms:
            f.write(t.to_bytes())
            f.write(struct.pack("<d", float(c)))


class _Streamer:
    def __init__(self, path: str):
        self._f = gzip.open(path, "rb")
        self._n, self._stride, self._size = struct.unpack("<QQQ", self._f.read(24))
        self._read = 0

    @classmethod
    def from_file(cls, path: str):
        return cls(path)

    def __iter__(self):
        return self

    def __next__(self):
        if self._read >= self._n:
            self._f.close()
            raise StopIteration
        kb = self._f.read(self._stride)
        (c,) = struct.unpack("<d", self._f.read(8))
        self._read += 1
        return self._decode(kb), c


class PauliTermStreamer(_Streamer):
    def _decode(self, kb: bytes) -> PauliString:
        half = len(kb) // 2
        return PauliString(int.from_bytes(kb[:half], "little"), int.from_bytes(kb[half:], "little"), self._size)


class MajoranaTermStreamer(_Streamer):
    def _decode(self, kb: bytes) -> MajoranaMonomial:
        return MajoranaMonomial.from_key(int.from_bytes(kb, "little"), self._size)
