"""At-rest operator types of the drop-in API.

``PauliString`` / ``MajoranaMonomial`` and the term sums mirror the reference's
PyO3 classes (``propaq/_rust_core/__init__.pyi``; Rust in
``crates/pauli/src/string.rs:30``, ``crates/pauli/src/termsum.rs:46``,
``crates/majorana/src/monomial.rs``, ``crates/majorana/src/termsum.rs``):
same constructor arguments, properties and algebra, so observables and
generators built for propaq work unchanged.  Term sums keep their terms as
arrays of packed u64 key words (the engine's interleaved layout) and FP64
coefficients, so results move between the device, files and the API without a
Python object per term.
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

def _words_for_bits(bits: int) -> int:
    return max(1, (bits + 63) // 64)


def _key_ints(keys: np.ndarray) -> list:
    """Engine keys (uint64[n, W]) as Python ints."""
    if keys.shape[1] == 1:
        return [int(k) for k in keys[:, 0]]
    return [sum(int(w) << (64 * i) for i, w in enumerate(row)) for row in keys]


def _int_words(key: int, W: int) -> np.ndarray:
    return np.array([(key >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(W)], dtype=np.uint64)


def _bits_of(words: np.ndarray) -> np.ndarray:
    """uint64[n, W] -> uint8[n, 64 W] of bits, least significant first."""
    b = np.ascontiguousarray(words, dtype="<u8").view(np.uint8).reshape(words.shape[0], 8 * words.shape[1])
    return np.unpackbits(b, axis=1, bitorder="little")


def _words_of(bits: np.ndarray, W: int) -> np.ndarray:
    """uint8[n, nbits] of bits -> uint64[n, W]."""
    n = bits.shape[0]
    full = np.zeros((n, 64 * W), dtype=np.uint8)
    full[:, :bits.shape[1]] = bits
    return np.packbits(full, axis=1, bitorder="little").view("<u8").astype(np.uint64).reshape(n, W)


_M1, _M2, _M4, _M8, _M16 = (np.uint64(v) for v in (0x3333333333333333, 0x0F0F0F0F0F0F0F0F, 0x00FF00FF00FF00FF,
                                                   0x0000FFFF0000FFFF, 0x00000000FFFFFFFF))
_EVEN_U = np.uint64(_EVEN)


def _gather_even(x: np.ndarray) -> np.ndarray:
    """The even bits of each u64 packed into its low 32 bits."""
    x = x & _EVEN_U
    x = (x | (x >> np.uint64(1))) & _M1
    x = (x | (x >> np.uint64(2))) & _M2
    x = (x | (x >> np.uint64(4))) & _M4
    x = (x | (x >> np.uint64(8))) & _M8
    return (x | (x >> np.uint64(16))) & _M16


def pauli_weights(keys: np.ndarray) -> np.ndarray:
    """Support of every key (crates/core/src/algebra/strings.rs:140)."""
    k = np.asarray(keys, dtype=np.uint64)
    return np.bitwise_count((k | (k >> np.uint64(1))) & _EVEN_U).sum(axis=1).astype(np.int64)


def majorana_weights(keys: np.ndarray, n_sites: int) -> np.ndarray:
    """Jordan-Wigner weight of every key (crates/majorana/src/algebra.rs:85-123):
    per site, single = x ^ y, occupied = x | y; the string bit is the prefix
    parity of singles (flipped when their total is odd); weight = popcount of
    single | (occupied ^ string) over the live sites."""
    k = np.asarray(keys, dtype=np.uint64)
    n, W = k.shape
    xs = [_gather_even(k[:, i]) for i in range(W)]
    ys = [_gather_even(k[:, i] >> np.uint64(1)) for i in range(W)]
    singles = [xs[i] ^ ys[i] for i in range(W)]
    total = sum(np.bitwise_count(s).astype(np.int64) for s in singles)
    flip = np.where(total & 1, np.uint64(0xFFFFFFFF), np.uint64(0))
    carry = np.zeros(n, dtype=np.uint64)
    w = np.zeros(n, dtype=np.int64)
    for i in range(W):
        lo = 32 * i
        live = 0 if lo >= n_sites else (0xFFFFFFFF if n_sites - lo >= 32 else (1 << (n_sites - lo)) - 1)
        live = np.uint64(live)
        p = singles[i]
        for sh in (1, 2, 4, 8, 16):
            p = (p ^ (p << np.uint64(sh))) & np.uint64(0xFFFFFFFF)
        p = p ^ carry
        carry = np.where((p >> np.uint64(31)) & np.uint64(1), np.uint64(0xFFFFFFFF), np.uint64(0))
        string = (p & live) ^ (flip & live)
        occ = xs[i] | ys[i]
        w += np.bitwise_count(singles[i] | (occ ^ string)).astype(np.int64)
    return w


class _TermSum:
    """A sum of basis strings with real coefficients, kept as arrays: keys in
    the engine's interleaved layout as uint64[n, W] (W words for the system
    size) and coefficients, in insertion order.  A key -> row index is built
    only when single terms are added or looked up; the engine hands its
    results over as arrays, and files are read and written a block at a time."""

    _term_cls = None  # set by subclasses

    def __init__(self, terms: Optional[dict] = None, dtype: Optional[str] = None):
        if dtype not in (None, "float64", "float32"):
            raise ValueError('dtype must be "float64" or "float32"')
        self._dtype = dtype or "float64"
        self._size = 0
        self._k = np.zeros((0, 1), dtype=np.uint64)
        self._c = np.zeros(0, dtype=np.float64)
        self._n = 0
        self._index: Optional[Dict[int, int]] = None
        if terms:
            for t, c in terms.items():
                self.add(t, c)

    # ---- storage
    def _key_bits(self, size: int) -> int:
        raise NotImplementedError

    def _W(self) -> int:
        return self._k.shape[1]

    def _widen(self, size: int) -> None:
        self._size = max(self._size, size)
        W = _words_for_bits(self._key_bits(self._size))
        if W > self._W():
            k = np.zeros((self._k.shape[0], W), dtype=np.uint64)
            k[:, :self._W()] = self._k
            self._k = k

    def _reserve(self, n: int) -> None:
        if n <= self._k.shape[0]:
            return
        cap = max(n, 2 * self._k.shape[0], 16)
        k = np.zeros((cap, self._W()), dtype=np.uint64)
        k[:self._n] = self._k[:self._n]
        c = np.zeros(cap, dtype=np.float64)
        c[:self._n] = self._c[:self._n]
        self._k, self._c = k, c

    def _lookup(self) -> Dict[int, int]:
        if self._index is None:
            self._index = {k: i for i, k in enumerate(_key_ints(self._k[:self._n]))}
        return self._index

    def _coerce(self, c: float) -> float:
        return float(np.float32(c)) if self._dtype == "float32" else float(c)

    def _coerce_array(self, c: np.ndarray) -> np.ndarray:
        return c.astype(np.float32).astype(np.float64) if self._dtype == "float32" else c

    def key_array(self, W: Optional[int] = None) -> np.ndarray:
        """The keys as uint64[n, W] (W at least the stored width)."""
        k = self._k[:self._n]
        if W is None or W == k.shape[1]:
            return k.copy()
        if W < k.shape[1]:
            if np.any(k[:, W:]):
                raise ValueError("keys do not fit the requested width")
            return k[:, :W].copy()
        out = np.zeros((self._n, W), dtype=np.uint64)
        out[:, :k.shape[1]] = k
        return out

    def coeff_array(self) -> np.ndarray:
        return self._c[:self._n].copy()

    # construction helper for the engine
    @classmethod
    def _from_arrays(cls, keys: np.ndarray, coeffs: np.ndarray, size: int) -> "_TermSum":
        out = cls()
        out._size = size
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        out._k = k.reshape(len(coeffs), -1) if k.size else np.zeros((0, _words_for_bits(out._key_bits(size))),
                                                                        dtype=np.uint64)
        out._widen(size)
        out._c = np.ascontiguousarray(coeffs, dtype=np.float64).copy()
        out._n = len(out._c)
        return out

    # ---- the reference's term-sum API
    @property
    def dtype(self) -> str:
        return self._dtype

    @property
    def n_units(self) -> int:
        return self._size

    def add(self, term, coeff: float) -> None:
        self._widen(self._term_size(term))
        k = term.key
        idx = self._lookup()
        r = idx.get(k)
        if r is not None:
            self._c[r] = self._coerce(self._c[r] + float(coeff))
            return
        self._reserve(self._n + 1)
        self._k[self._n] = _int_words(k, self._W())
        self._c[self._n] = self._coerce(float(coeff))
        idx[k] = self._n
        self._n += 1

    def __setitem__(self, term, coeff: float) -> None:
        self._widen(self._term_size(term))
        k = term.key
        idx = self._lookup()
        r = idx.get(k)
        if r is None:
            self.add(term, coeff)
        else:
            self._c[r] = self._coerce(coeff)

    def __getitem__(self, term) -> float:
        r = self._lookup().get(term.key)
        return float(self._c[r]) if r is not None else 0.0

    def scale(self, factor: float) -> None:
        self._c[:self._n] = self._coerce_array(self._c[:self._n] * float(factor))

    def merge(self, other: "_TermSum") -> None:
        """Add other's terms in other's order, as add() of each would: a key
        already present gets its coefficients summed in that order, new keys
        follow in order of first appearance."""
        if other._dtype != self._dtype:
            raise ValueError("both term sums must share the same dtype")
        self._widen(other._size)
        W = self._W()
        allk = np.vstack([self._k[:self._n], other.key_array(W)])
        allc = np.concatenate([self._c[:self._n], other._c[:other._n]])
        view = np.ascontiguousarray(allk).view(np.dtype((np.void, 8 * W))).reshape(-1)
        _, first, inv = np.unique(view, return_index=True, return_inverse=True)
        inv = inv.reshape(-1)
        res = allc[first].copy()
        rest = np.ones(len(allc), dtype=bool)
        rest[first] = False
        rest = np.flatnonzero(rest)
        if len(rest):
            if self._dtype == "float32":
                for i in rest:                      # rounded after every add, as add() does
                    res[inv[i]] = float(np.float32(res[inv[i]] + allc[i]))
            else:
                np.add.at(res, inv[rest], allc[rest])   # unbuffered, in index order
        order = np.argsort(first, kind="stable")
        self._k = allk[first[order]]
        self._c = self._coerce_array(res[order])
        self._n = len(self._c)
        self._index = None

    def norm_squared(self) -> float:
        c = self._c[:self._n]
        return float(np.dot(c, c))

    def weights(self) -> np.ndarray:
        """The weight of every term (Pauli support, Majorana Jordan-Wigner weight)."""
        raise NotImplementedError

    def items(self):
        return [(self._make(k), float(c)) for k, c in zip(_key_ints(self._k[:self._n]), self._c[:self._n])]

    def keys_and_coeffs(self) -> Tuple[list, list]:
        return _key_ints(self._k[:self._n]), [float(c) for c in self._c[:self._n]]

    @property
    def _terms(self) -> Dict[int, float]:
        """key -> coefficient as a dict (a copy)."""
        return dict(zip(_key_ints(self._k[:self._n]), (float(c) for c in self._c[:self._n])))

    def __len__(self) -> int:
        return self._n

    def __iter__(self) -> Iterator:
        return iter(self.items())

    def copy(self) -> "_TermSum":
        out = type(self)(dtype=self._dtype)
        out._size = self._size
        out._k = self._k[:self._n].copy()
        out._c = self._c[:self._n].copy()
        out._n = self._n
        return out

    @property
    def sparse_key_bytes(self) -> int:
        return self._n * 8 * _words_for_bits(2 * self._units())

    def _keep(self, mask: np.ndarray) -> None:
        self._k = self._k[:self._n][mask]
        self._c = self._c[:self._n][mask]
        self._n = len(self._c)
        self._index = None

    def truncate(self, policy) -> None:
        """Drop terms the policy rejects (TruncationPolicy.should_truncate)."""
        w = self.weights()
        mags = np.abs(self._c[:self._n])
        from .truncation import TruncationPolicy
        if type(policy) is TruncationPolicy:
            drop = mags < float(policy.coeff_cutoff)
            if policy.weight_cutoff is not None:
                drop |= w > int(policy.weight_cutoff)
        else:
            drop = np.array([bool(policy.should_truncate(int(x), float(m))) for x, m in zip(w, mags)], dtype=bool)
        self._keep(~drop)

    def apply_damping(self, noise, active_modes: int) -> None:
        w = self.weights()
        table = {int(x): float(noise.damping_factor(int(x), active_modes)) for x in np.unique(w)}
        f = np.array([table[int(x)] for x in w], dtype=np.float64)
        self._c[:self._n] = self._coerce_array(self._c[:self._n] * f)

    # gzip term file (crates/core/src/interface/term_io.rs:13-89)
    def _key_bytes(self) -> Tuple[np.ndarray, int]:
        """Every key in the file's byte layout (uint8[n, stride]) and the stride."""
        raise NotImplementedError

    def save(self, path: str) -> None:
        with gzip.open(path, "wb") as f:
            if not self._n:
                f.write(struct.pack("<QQQ", 0, 0, 0))
                return
            kb, stride = self._key_bytes()
            f.write(struct.pack("<QQQ", self._n, stride, self._size))
            if self._n:
                rec = np.zeros((self._n, stride + 8), dtype=np.uint8)
                rec[:, :stride] = kb
                rec[:, stride:] = self._c[:self._n].astype("<f8").view(np.uint8).reshape(self._n, 8)
                f.write(rec.tobytes())

    @classmethod
    def from_file(cls, path: str) -> "_TermSum":
        with gzip.open(path, "rb") as f:
            n, stride, size = struct.unpack("<QQQ", f.read(24))
            data = np.frombuffer(f.read(n * (stride + 8)), dtype=np.uint8).reshape(n, stride + 8)
        out = cls()
        if n == 0:
            return out
        keys = cls._keys_from_bytes(data[:, :stride], stride, size)
        coeffs = data[:, stride:].copy().view("<f8").reshape(n).astype(np.float64)
        out._size = size
        out._widen(size)
        out.merge(cls._from_arrays(keys, coeffs, size))   # folds repeated keys as add() would
        return out

    def merge_from_file(self, streamer) -> None:
        for t, c in streamer:
            self.add(t, c)


class PauliTermSum(_TermSum):
    """A weighted sum of Pauli strings with real coefficients."""

    def _key_bits(self, size: int) -> int:
        return 2 * size

    def _term_size(self, t: PauliString) -> int:
        return t.n_qubits

    def _units(self) -> int:
        return self._size

    def _make(self, key: int) -> PauliString:
        return PauliString.from_key(key, self._size)

    def weights(self) -> np.ndarray:
        return pauli_weights(self._k[:self._n])

    def _key_bytes(self):
        nb = max(1, (self._size + 7) // 8)
        bits = _bits_of(self._k[:self._n])
        x = np.zeros((self._n, 8 * nb), dtype=np.uint8)
        z = np.zeros((self._n, 8 * nb), dtype=np.uint8)
        m = min(self._size, bits.shape[1] // 2)
        x[:, :m] = bits[:, 0:2 * m:2]
        z[:, :m] = bits[:, 1:2 * m:2]
        return np.hstack([np.packbits(x, axis=1, bitorder="little"),
                          np.packbits(z, axis=1, bitorder="little")]), 2 * nb

    @classmethod
    def _keys_from_bytes(cls, kb: np.ndarray, stride: int, size: int) -> np.ndarray:
        nb = stride // 2
        x = np.unpackbits(kb[:, :nb], axis=1, bitorder="little")[:, :size]
        z = np.unpackbits(kb[:, nb:], axis=1, bitorder="little")[:, :size]
        bits = np.zeros((kb.shape[0], 2 * size), dtype=np.uint8)
        bits[:, 0::2] = x
        bits[:, 1::2] = z
        return _words_of(bits, _words_for_bits(2 * size))

    @classmethod
    def _load(cls, path):
        return PauliTermStreamer.from_file(path)


class MajoranaTermSum(_TermSum):
    """A weighted sum of Majorana monomials with real coefficients."""

    def _key_bits(self, size: int) -> int:
        return size

    def _term_size(self, t: MajoranaMonomial) -> int:
        return t.n_modes

    def _units(self) -> int:
        return self._size // 2

    def _make(self, key: int) -> MajoranaMonomial:
        return MajoranaMonomial.from_key(key, self._size)

    def weights(self) -> np.ndarray:
        return majorana_weights(self._k[:self._n], self._size // 2)

    def _key_bytes(self):
        nb = max(1, (self._size + 7) // 8)
        raw = np.ascontiguousarray(self._k[:self._n], dtype="<u8").view(np.uint8).reshape(self._n, 8 * self._W())
        out = np.zeros((self._n, nb), dtype=np.uint8)
        m = min(nb, raw.shape[1])
        out[:, :m] = raw[:, :m]
        return out, nb

    @classmethod
    def _keys_from_bytes(cls, kb: np.ndarray, stride: int, size: int) -> np.ndarray:
        W = _words_for_bits(size)
        raw = np.zeros((kb.shape[0], 8 * W), dtype=np.uint8)
        raw[:, :min(stride, 8 * W)] = kb[:, :min(stride, 8 * W)]
        return raw.view("<u8").astype(np.uint64).reshape(kb.shape[0], W)

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
