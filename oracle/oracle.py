"""TEST INFRASTRUCTURE ONLY: Python face of the oracle.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use
this module, and only as the checker (or, for the baseline, as the timed CPU
port). The product library never imports it.

* ``Plan`` wraps oracle/liboracle.so, the plain-C restatement of the reference's
  RNS + negacyclic NTT + CRT (polymult_oracle.c names the reference lines each
  step follows).
* ``RefSolver`` wraps oracle/_ref/libref_solve_cpu.so, the reference's own
  solve_cpu.cpp compiled from its sources (present only where it was built).
* The pure-Python helpers below restate the arithmetic a second time, with Python
  integers, for tiny cases: the negacyclic schoolbook product, the NTT by its
  definition, and the closed forms of the reference's adversarial patterns
  (fuzz.py:53-63).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_solve_cpu.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib = None


def build(ref: bool = True) -> None:
    """Compiles liboracle.so (and _ref when the reference sources are present)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = ctypes.CDLL(LIB)
        L.pmo_plan_new.restype = ctypes.c_void_p
        L.pmo_plan_new.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.pmo_plan_free.argtypes = [ctypes.c_void_p]
        L.pmo_plan_K.argtypes = [ctypes.c_void_p]
        L.pmo_plan_K.restype = ctypes.c_int
        for name in ("pmo_plan_primes", "pmo_plan_psi", "pmo_plan_M", "pmo_plan_negvP", "pmo_plan_scale"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = _u32p
        L.pmo_choose_primes.argtypes = [ctypes.c_uint32, ctypes.c_int, _u32p, ctypes.c_int]
        L.pmo_choose_primes.restype = ctypes.c_int
        L.pmo_decompose.argtypes = [ctypes.c_void_p, _u32p, _u32p]
        L.pmo_ntt_forward.argtypes = [ctypes.c_void_p, _u32p]
        L.pmo_ntt_inverse.argtypes = [ctypes.c_void_p, _u32p]
        L.pmo_pointwise.argtypes = [ctypes.c_void_p, _u32p, _u32p]
        L.pmo_crt.argtypes = [ctypes.c_void_p, _u32p, _u32p]
        L.pmo_crt.restype = ctypes.c_double
        L.pmo_multiply.argtypes = [ctypes.c_void_p, _u32p, _u32p, _u32p]
        L.pmo_multiply.restype = ctypes.c_double
        L.pmo_set_threads.argtypes = [ctypes.c_int]
        L.pmo_max_threads.restype = ctypes.c_int
        L.pmo_gen_q.argtypes = [ctypes.c_uint64, ctypes.c_uint32, _u32p]
        L.pmo_gen_operand.argtypes = [ctypes.c_uint64, ctypes.c_uint32, _u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        L.pmo_gen_operand.restype = ctypes.c_int
        L.pmo_schoolbook.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_int, _u32p, _u32p, _u32p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def set_threads(n: int) -> None:
    lib().pmo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().pmo_max_threads())


# --------------------------------------------------------------------------
# Limbs <-> Python integers

def limbs_of(x: int, L: int) -> list:
    return [(x >> (32 * i)) & 0xFFFFFFFF for i in range(L)]


def int_of(limbs) -> int:
    return sum(int(v) << (32 * i) for i, v in enumerate(limbs))


def to_array(values, L: int) -> np.ndarray:
    """Python integers -> (len, L) uint32 limbs."""
    out = np.zeros((len(values), L), dtype=np.uint32)
    for i, x in enumerate(values):
        out[i] = np.frombuffer(int(x).to_bytes(4 * L, "little"), dtype="<u4")
    return out


def to_ints(arr: np.ndarray) -> list:
    arr = np.ascontiguousarray(arr, dtype=np.uint32)
    L = arr.shape[-1]
    raw = arr.reshape(-1, L).astype("<u4").tobytes()
    return [int.from_bytes(raw[4 * L * i:4 * L * (i + 1)], "little") for i in range(arr.reshape(-1, L).shape[0])]


def digest(arr: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr, dtype="<u4").tobytes()).hexdigest()


# --------------------------------------------------------------------------
# The C restatement

class Plan:
    """The oracle's plan for one point (N, q): rns_plan.h:118-249 restated."""

    def __init__(self, N: int, q_limbs: np.ndarray):
        q_limbs = np.ascontiguousarray(q_limbs, dtype=np.uint32)
        err = ctypes.create_string_buffer(256)
        self.N, self.L = int(N), int(q_limbs.shape[0])
        self._q = q_limbs
        self._h = lib().pmo_plan_new(self.N, _ptr(q_limbs), self.L, err, 256)
        if not self._h:
            raise ValueError(err.value.decode())
        self.K = int(lib().pmo_plan_K(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pmo_plan_free(self._h)
            self._h = None

    def _vec(self, fn, n):
        return np.ctypeslib.as_array(getattr(lib(), fn)(self._h), shape=(n,)).copy()

    @property
    def primes(self) -> np.ndarray:
        return self._vec("pmo_plan_primes", self.K)

    @property
    def psi(self) -> np.ndarray:
        return self._vec("pmo_plan_psi", self.K)

    @property
    def scale(self) -> np.ndarray:
        return self._vec("pmo_plan_scale", self.K)

    @property
    def M(self) -> np.ndarray:
        return self._vec("pmo_plan_M", self.K * self.L).reshape(self.K, self.L)

    @property
    def negvP(self) -> np.ndarray:
        return self._vec("pmo_plan_negvP", (self.K + 1) * self.L).reshape(self.K + 1, self.L)

    def decompose(self, limbs: np.ndarray) -> np.ndarray:
        limbs = np.ascontiguousarray(limbs, dtype=np.uint32)
        assert limbs.shape == (self.N, self.L)
        res = np.empty((self.K, self.N), dtype=np.uint32)
        lib().pmo_decompose(self._h, _ptr(limbs), _ptr(res))
        return res

    def ntt_forward(self, res: np.ndarray) -> np.ndarray:
        out = np.array(res, dtype=np.uint32, order="C", copy=True)
        assert out.shape == (self.K, self.N)
        lib().pmo_ntt_forward(self._h, _ptr(out))
        return out

    def ntt_inverse(self, res: np.ndarray) -> np.ndarray:
        out = np.array(res, dtype=np.uint32, order="C", copy=True)
        assert out.shape == (self.K, self.N)
        lib().pmo_ntt_inverse(self._h, _ptr(out))
        return out

    def pointwise(self, A: np.ndarray, B: np.ndarray) -> np.ndarray:
        out = np.array(A, dtype=np.uint32, order="C", copy=True)
        B = np.ascontiguousarray(B, dtype=np.uint32)
        lib().pmo_pointwise(self._h, _ptr(out), _ptr(B))
        return out

    def crt(self, Y: np.ndarray):
        """(c, worst distance of sum y_j / p_j from an integer)."""
        Y = np.ascontiguousarray(Y, dtype=np.uint32)
        c = np.empty((self.N, self.L), dtype=np.uint32)
        worst = lib().pmo_crt(self._h, _ptr(Y), _ptr(c))
        return c, float(worst)

    def multiply(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        assert a.shape == (self.N, self.L) and b.shape == (self.N, self.L)
        c = np.empty((self.N, self.L), dtype=np.uint32)
        self.last_margin = lib().pmo_multiply(self._h, _ptr(a), _ptr(b), _ptr(c))
        return c


def choose_primes(N: int, qbits: int) -> np.ndarray:
    buf = np.zeros(200, dtype=np.uint32)
    K = lib().pmo_choose_primes(int(N), int(qbits), _ptr(buf), 200)
    if K < 0:
        raise ValueError("no prime set (code %d)" % K)
    return buf[:K].copy()


def gen_q(seed: int, W: int) -> np.ndarray:
    q = np.zeros((W + 31) // 32, dtype=np.uint32)
    lib().pmo_gen_q(int(seed), int(W), _ptr(q))
    return q


def gen_operand(seed: int, stream: int, q: np.ndarray, W: int, N: int) -> np.ndarray:
    L = (W + 31) // 32
    out = np.empty((N, L), dtype=np.uint32)
    q = np.ascontiguousarray(q, dtype=np.uint32)
    if lib().pmo_gen_operand(int(seed), int(stream), _ptr(q), int(W), int(N), _ptr(out)) != 0:
        raise RuntimeError("generator exhausted its 64 attempts")
    return out


def gen_point(seed: int, N: int, W: int):
    """(q, a, b) for a seeded random point: q from stream 0, a stream 1, b stream 2."""
    q = gen_q(seed, W)
    return q, gen_operand(seed, 1, q, W, N), gen_operand(seed, 2, q, W, N)


def schoolbook_c(N: int, q: np.ndarray, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.uint32)
    L = q.shape[0]
    c = np.empty((N, L), dtype=np.uint32)
    lib().pmo_schoolbook(int(N), _ptr(q), L, _ptr(np.ascontiguousarray(a, dtype=np.uint32)),
                         _ptr(np.ascontiguousarray(b, dtype=np.uint32)), _ptr(c))
    return c


# --------------------------------------------------------------------------
# The reference's own CPU path (oracle/_ref), when built

def ref_available() -> bool:
    return os.path.exists(REF_LIB)


class RefSolver:
    """fherma_init / fherma_run / fherma_free of the reference's solve_cpu.cpp."""

    _lib = None

    def __init__(self, N: int, W: int, q: np.ndarray):
        if RefSolver._lib is None:
            L = ctypes.CDLL(REF_LIB)
            L.ref_init.restype = ctypes.c_void_p
            L.ref_init.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p]
            L.ref_run.restype = ctypes.c_int
            L.ref_run.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, _u32p]
            L.ref_free.argtypes = [ctypes.c_void_p]
            L.ref_last_error.restype = ctypes.c_char_p
            RefSolver._lib = L
        q = np.ascontiguousarray(q, dtype=np.uint32)
        self.N, self.W, self.L = int(N), int(W), int(q.shape[0])
        self._h = RefSolver._lib.ref_init(self.N, self.W, self.L, _ptr(q))
        if not self._h:
            raise ValueError(RefSolver._lib.ref_last_error().decode())

    def run(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        c = np.empty((self.N, self.L), dtype=np.uint32)
        rc = RefSolver._lib.ref_run(self._h, self.N, self.L, _ptr(np.ascontiguousarray(a, dtype=np.uint32)),
                                    _ptr(np.ascontiguousarray(b, dtype=np.uint32)), _ptr(c))
        if rc != 0:
            raise RuntimeError(RefSolver._lib.ref_last_error().decode())
        return c

    def close(self):
        if self._h:
            RefSolver._lib.ref_free(self._h)
            self._h = None

    def __del__(self):
        if getattr(self, "_h", None):
            self.close()


# --------------------------------------------------------------------------
# Pure-Python second restatement for tiny cases

def py_negacyclic(a: list, b: list, q: int) -> list:
    """Exact O(N^2) negacyclic product over the integers, then mod q."""
    N = len(a)
    c = [0] * N
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                k = i + j
                if k < N:
                    c[k] += x * y
                else:
                    c[k - N] -= x * y
    return [v % q for v in c]


def py_ntt_definition(x: list, p: int, psi: int) -> list:
    """X[k] = sum_i x_i psi^((2 brv(k) + 1) i) mod p: the bit-reversed-order
    negacyclic NTT that the merged Cooley-Tukey transform computes."""
    N = len(x)
    logN = N.bit_length() - 1
    out = []
    for k in range(N):
        e = 2 * int(format(k, "0%db" % logN)[::-1], 2) + 1 if logN else 1
        w = pow(psi, e, p)
        out.append(sum(xi * pow(w, i, p) for i, xi in enumerate(x)) % p)
    return out


def patterns(N: int, q: int, rand_a: list, rand_b: list) -> dict:
    """The reference's adversarial operand pairs (fuzz.py:53-63), each with the
    answer its closed form gives: name -> (a, b, c)."""
    qm = q - 1
    out = {}
    out["all_q_minus_1"] = ([qm] * N, [qm] * N, [(2 * k + 2 - N) % q for k in range(N)])
    out["zero"] = ([0] * N, list(rand_b), [0] * N)
    wrap_c = [0] * N
    if N >= 2:
        wrap_c[N - 2] = 1 % q          # x^(N-1) * (q-1) x^(N-1) = -(q-1) x^(N-2)
    else:
        wrap_c[0] = 1 % q
    out["wrap"] = ([0] * (N - 1) + [1 % q], [0] * (N - 1) + [qm], wrap_c)
    out["identity"] = (list(rand_a), [1 % q] + [0] * (N - 1), [x % q for x in rand_a])
    alt_c = [0 if k % 2 == 0 else (k + 1 - N // 2) % q for k in range(N)]
    out["alternating"] = ([qm * (i % 2) for i in range(N)], [qm * ((i + 1) % 2) for i in range(N)], alt_c)
    return out
