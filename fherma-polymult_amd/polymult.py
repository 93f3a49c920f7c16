"""Python face of libfherma_polymult.so (ctypes over include/fherma_polymult.h).

The product path is the HIP library: there is no CPU fallback. If the built
library is missing, importing this module raises, naming the build command.

    plan = Plan(N, W, q)               # fherma_init      (reference solve.cu:710)
    c = plan.multiply(a, b)            # fherma_run       (reference solve.cu:871)
    plan.close()                       # fherma_free      (reference solve.cu:1061)

Operands are numpy uint32 arrays of shape (N, L): little-endian limbs,
coefficient-major, as the spec's wire format (reference fherma.h:34-41).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfherma_polymult.so")

PM_OK, PM_EINVAL, PM_EHIP, PM_ENOMEM, PM_EINTERNAL = 0, 1, 2, 3, 4
FORM_PLAIN, FORM_MFMA = 0, 1
HOST_STAGED, HOST_MAPPED, HOST_DIRECT = 0, 1, 2
TWIDDLES_TABLE, TWIDDLES_GEN = 0, 1
NTT_VALU, NTT_MFMA = 0, 1
NPHASES = 7

_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


class BatchShape(ctypes.Structure):
    """pm_batch_shape: runs, streams, ntt_form, twiddles."""
    _fields_ = [("runs", ctypes.c_int), ("streams", ctypes.c_int), ("ntt_form", ctypes.c_int),
                ("twiddles", ctypes.c_int)]


class PolymulError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError("libfherma_polymult.so is not built: run `python fherma-polymult_amd/build.py` "
                          "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "pm_last_error": (ctypes.c_char_p, []),
        "pm_plan_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                          ctypes.POINTER(_vp)]),
        "pm_plan_destroy": (None, [_vp]),
        "pm_plan_query": (ctypes.c_int, [_vp, _u32p, _u32p, _u32p]),
        "pm_plan_primes": (ctypes.c_int, [_vp, _u32p, ctypes.c_uint32]),
        "pm_plan_set_form": (ctypes.c_int, [_vp, ctypes.c_int]),
        "pm_choose_primes": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p]),
        "pm_multiply": (ctypes.c_int, [_vp, _u32p, _u32p, _u32p]),
        "pm_plan_set_host_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
        "pm_plan_set_host_modes": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
        "pm_plan_get_host_modes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "pm_plan_tune_host": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
        "pm_plan_tune_kernels": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
        "pm_multiply_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "pm_multiply_device_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "pm_batch_multiply_device": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
        "pm_batch_multiply_host": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                  ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int]),
        "pm_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                           ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
        "pm_batch_create_shaped": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                  ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                  ctypes.POINTER(BatchShape), ctypes.POINTER(_vp)]),
        "pm_batch_get_shape": (ctypes.c_int, [_vp, ctypes.POINTER(BatchShape), ctypes.POINTER(ctypes.c_double)]),
        "pm_batch_launch": (ctypes.c_int, [_vp, _vp]),
        "pm_batch_phased": (ctypes.c_int, [_vp]),
        "pm_plan_set_twiddles": (ctypes.c_int, [_vp, ctypes.c_int]),
        "pm_plan_set_ntt_form": (ctypes.c_int, [_vp, ctypes.c_int]),
        "pm_plan_get_ntt_form": (ctypes.c_int, [_vp]),
        "pm_plan_set_fused_cols": (ctypes.c_int, [_vp, ctypes.c_int]),
        "pm_plan_get_fused_cols": (ctypes.c_int, [_vp]),
        "pm_plan_get_twiddles": (ctypes.c_int, [_vp]),
        "pm_batch_destroy": (None, [_vp]),
        "pm_stage_decompose": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
        "pm_stage_ntt": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
        "pm_stage_pointwise": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
        "pm_stage_crt": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
        "pm_time_multiply_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_double), _vp]),
        "pm_profile_phases": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                             _vp]),
        "pm_phase_name": (ctypes.c_char_p, [ctypes.c_int]),
        "pm_device_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
        "pm_device_free": (ctypes.c_int, [_vp]),
        "pm_copy_to_device": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
        "pm_copy_to_host": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
        "pm_device_copy": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
        "pm_device_sum": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)]),
        "pm_stream_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
        "pm_stream_destroy": (ctypes.c_int, [_vp]),
        "pm_stream_synchronize": (ctypes.c_int, [_vp]),
        "pm_device_synchronize": (ctypes.c_int, []),
        "pm_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]),
        "pm_probe_issue": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 4),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


lib = _load()


def _check(rc: int) -> None:
    if rc != PM_OK:
        raise PolymulError(rc, lib.pm_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    if a.dtype != np.uint32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous uint32 array")
    return a.ctypes.data_as(_u32p)


def choose_primes(N: int, qbits: int) -> np.ndarray:
    """The plan's prime rule on the host alone (no device needed): the
    reference's primes below 2^31 (rns_plan.h:138-155)."""
    buf = np.zeros(256, dtype=np.uint32)
    K = ctypes.c_uint32(0)
    _check(lib.pm_choose_primes(int(N), int(qbits), _ptr(buf), 256, ctypes.byref(K)))
    return buf[:K.value].copy()


# --------------------------------------------------------------------------
# Device memory

class DeviceBuffer:
    """A device allocation of `nbytes`, freed on close()."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        _check(lib.pm_device_alloc(self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array of %d bytes does not fit a buffer of %d" % (arr.nbytes, self.nbytes))
        _check(lib.pm_copy_to_device(_vp(self.ptr), arr.ctypes.data_as(_vp), arr.nbytes))

    def download(self, shape, dtype=np.uint32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("%d bytes requested from a buffer of %d" % (out.nbytes, self.nbytes))
        _check(lib.pm_copy_to_host(out.ctypes.data_as(_vp), _vp(self.ptr), out.nbytes))
        return out

    def close(self) -> None:
        if self.ptr:
            _check(lib.pm_device_free(_vp(self.ptr)))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    """A HIP stream from the library (pm_stream_create); pass it as `stream=`."""

    def __init__(self):
        h = _vp()
        _check(lib.pm_stream_create(ctypes.byref(h)))
        self.handle = h.value

    def synchronize(self) -> None:
        _check(lib.pm_stream_synchronize(_vp(self.handle)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.pm_stream_destroy(_vp(self.handle))
            self.handle = None

    def __del__(self):
        self.close()


def _fit(buf: DeviceBuffer, nbytes: int, what: str) -> ctypes.c_void_p:
    """buf's device pointer, after checking that it holds `nbytes`: the kernels
    trust the plan's sizes, so a short buffer is refused here, on the host."""
    if not buf.ptr:
        raise ValueError("%s: buffer is closed" % what)
    if buf.nbytes < nbytes:
        raise ValueError("%s: buffer of %d bytes, the plan needs %d" % (what, buf.nbytes, nbytes))
    return _vp(buf.ptr)


def _sv(stream):
    """A stream argument: None (the default stream or the plan's), a Stream, or a raw handle."""
    return _vp(stream.handle if isinstance(stream, Stream) else stream)


def synchronize() -> None:
    _check(lib.pm_device_synchronize())


def device_copy(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int = None, stream=None) -> None:
    """Device-to-device copy by the library's plain streaming kernel."""
    n = min(dst.nbytes, src.nbytes) if nbytes is None else int(nbytes)
    if n > dst.nbytes or n > src.nbytes:
        raise ValueError("copy larger than its buffers")
    _check(lib.pm_device_copy(_vp(dst.ptr), _vp(src.ptr), n, _sv(stream)))


def device_sum(src: DeviceBuffer, nbytes: int = None) -> int:
    """Sum mod 2^64 of the buffer's 32-bit words, by a read-only kernel (synchronous)."""
    n = src.nbytes if nbytes is None else int(nbytes)
    if n > src.nbytes:
        raise ValueError("sum larger than its buffer")
    out = ctypes.c_uint64()
    _check(lib.pm_device_sum(_vp(src.ptr), n, ctypes.byref(out)))
    return out.value


def device_info():
    name = ctypes.create_string_buffer(128)
    cus = ctypes.c_int(0)
    _check(lib.pm_device_info(name, 128, ctypes.byref(cus)))
    return name.value.decode(), cus.value


PROBE_KINDS = ["v_add_u32", "v_mul_lo_u32", "v_mul_hi_u32", "v_mad_u64_u32", "v_mad_i64_i32",
               "v_permlane32_swap_b32", "v_permlane16_swap_b32", "v_mfma_i32_16x16x64_i8",
               "mfma_waves_beside_mul_lo_waves", "mfma_and_mul_lo_in_one_wave", "mfma_waves_beside_add_waves",
               "mfma_and_add_in_one_wave"]


def probe_issue(kind: int, iters: int) -> dict:
    """pm_probe_issue: wall time of an issue-rate probe and what each SIMD issued."""
    ms, valu, mfma, mhz = (ctypes.c_double() for _ in range(4))
    _check(lib.pm_probe_issue(int(kind), int(iters), ctypes.byref(ms), ctypes.byref(valu), ctypes.byref(mfma),
                              ctypes.byref(mhz)))
    return {"kind": PROBE_KINDS[kind], "ms": ms.value, "valu_per_simd": valu.value, "mfma_per_simd": mfma.value,
            "clock_mhz": mhz.value}


# --------------------------------------------------------------------------
# Plans

class Plan:
    """One point (N, W, q): primes, tables and workspace on the device."""

    def __init__(self, N: int, W: int, q: np.ndarray, form: int | None = None):
        q = np.ascontiguousarray(q, dtype=np.uint32)
        h = _vp()
        _check(lib.pm_plan_create(int(N), int(W), int(q.shape[0]), _ptr(q), ctypes.byref(h)))
        self._h = h.value
        n, l, k = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib.pm_plan_query(_vp(self._h), ctypes.byref(n), ctypes.byref(l), ctypes.byref(k)))
        self.N, self.L, self.K, self.W = n.value, l.value, k.value, int(W)
        self.form = None
        if form is not None:
            self.set_form(form)

    @property
    def handle(self):
        return _vp(self._h)

    def set_form(self, form: int) -> None:
        _check(lib.pm_plan_set_form(self.handle, int(form)))
        self.form = form

    @property
    def primes(self) -> np.ndarray:
        out = np.zeros(self.K, dtype=np.uint32)
        _check(lib.pm_plan_primes(self.handle, _ptr(out), self.K))
        return out

    def _operand(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.uint32)
        if x.shape != (self.N, self.L):
            raise ValueError("operand must have shape (N, L) = (%d, %d)" % (self.N, self.L))
        return x

    def multiply(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a, b = self._operand(a), self._operand(b)
        c = np.empty((self.N, self.L), dtype=np.uint32)
        _check(lib.pm_multiply(self.handle, _ptr(a), _ptr(b), _ptr(c)))
        return c

    def set_ntt_form(self, form: int) -> None:
        """NTT_VALU or NTT_MFMA: the column passes' butterflies or matrix-core form."""
        _check(lib.pm_plan_set_ntt_form(self.handle, int(form)))

    @property
    def ntt_form(self) -> int:
        return lib.pm_plan_get_ntt_form(self.handle)

    def set_fused_cols(self, on: bool) -> None:
        """Run the first forward column stages inside decompose and the row pass
        (N = 65536, matrix-core form), or the separate column passes."""
        _check(lib.pm_plan_set_fused_cols(self.handle, 1 if on else 0))

    @property
    def fused_cols(self) -> int:
        return lib.pm_plan_get_fused_cols(self.handle)

    def set_twiddles(self, source: int) -> None:
        """Row-pass twiddle source: TWIDDLES_TABLE or TWIDDLES_GEN (same results)."""
        _check(lib.pm_plan_set_twiddles(self.handle, int(source)))

    @property
    def twiddles(self) -> int:
        return lib.pm_plan_get_twiddles(self.handle)

    def set_host_mode(self, mode: int) -> None:
        _check(lib.pm_plan_set_host_mode(self.handle, int(mode)))

    def set_host_modes(self, in_mode: int, out_mode: int) -> None:
        """The input side's and the output side's host mode separately."""
        _check(lib.pm_plan_set_host_modes(self.handle, int(in_mode), int(out_mode)))

    @property
    def host_modes(self):
        i, o = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.pm_plan_get_host_modes(self.handle, ctypes.byref(i), ctypes.byref(o)))
        return i.value, o.value

    def tune_host(self):
        """(chosen host mode, its median ms per product) from the timed choice."""
        m, ms = ctypes.c_int(0), ctypes.c_double(0)
        _check(lib.pm_plan_tune_host(self.handle, ctypes.byref(m), ctypes.byref(ms)))
        return m.value, ms.value

    def tune_kernels(self) -> float:
        """The timed choice of the row-pass twiddle source and the column-pass
        form; returns the chosen setting's median ms per device product."""
        ms = ctypes.c_double(0)
        _check(lib.pm_plan_tune_kernels(self.handle, ctypes.byref(ms)))
        return ms.value

    @property
    def limb_bytes(self) -> int:
        """Bytes of one operand or product: N x L words."""
        return 4 * self.N * self.L

    @property
    def residue_bytes(self) -> int:
        """Bytes of one residue array: K x N words."""
        return 4 * self.K * self.N

    def _abc(self, a: DeviceBuffer, b: DeviceBuffer, c: DeviceBuffer):
        n = self.limb_bytes
        return _fit(a, n, "a"), _fit(b, n, "b"), _fit(c, n, "c")

    def multiply_device(self, a: DeviceBuffer, b: DeviceBuffer, c: DeviceBuffer, stream=None) -> None:
        _check(lib.pm_multiply_device(self.handle, *self._abc(a, b, c), _sv(stream)))

    def multiply_device_noalloc(self, a: DeviceBuffer, b: DeviceBuffer, c: DeviceBuffer, stream=None) -> None:
        """multiply_device with the residues in the library's static workspace:
        no allocation in the call (pm_multiply_device_noalloc)."""
        _check(lib.pm_multiply_device_noalloc(self.handle, *self._abc(a, b, c), _sv(stream)))

    # the five steps, device to device
    def stage_decompose(self, limbs: DeviceBuffer, res: DeviceBuffer, stream=None) -> None:
        _check(lib.pm_stage_decompose(self.handle, _fit(limbs, self.limb_bytes, "limbs"),
                                      _fit(res, self.residue_bytes, "res"), _sv(stream)))

    def stage_ntt(self, res: DeviceBuffer, forward: bool, stream=None) -> None:
        _check(lib.pm_stage_ntt(self.handle, _fit(res, self.residue_bytes, "res"), 1 if forward else 0, _sv(stream)))

    def stage_pointwise(self, A: DeviceBuffer, B: DeviceBuffer, stream=None) -> None:
        _check(lib.pm_stage_pointwise(self.handle, _fit(A, self.residue_bytes, "A"), _fit(B, self.residue_bytes, "B"),
                                      _sv(stream)))

    def stage_crt(self, res: DeviceBuffer, c: DeviceBuffer, stream=None) -> None:
        _check(lib.pm_stage_crt(self.handle, _fit(res, self.residue_bytes, "res"), _fit(c, self.limb_bytes, "c"),
                                _sv(stream)))

    def time_device(self, a: DeviceBuffer, b: DeviceBuffer, c: DeviceBuffer, warmup: int, iters: int,
                    stream=None) -> float:
        ms = ctypes.c_double(0)
        _check(lib.pm_time_multiply_device(self.handle, *self._abc(a, b, c), int(warmup), int(iters),
                                           ctypes.byref(ms), _sv(stream)))
        return ms.value

    def profile_phases(self, a: DeviceBuffer, b: DeviceBuffer, c: DeviceBuffer, iters: int, stream=None) -> dict:
        out = (ctypes.c_double * NPHASES)()
        _check(lib.pm_profile_phases(self.handle, *self._abc(a, b, c), int(iters), out, _sv(stream)))
        return {lib.pm_phase_name(i).decode(): out[i] for i in range(NPHASES)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.pm_plan_destroy(_vp(self._h))
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _batch_args(plans, a_bufs, b_bufs, c_bufs):
    n = len(plans)
    if not (len(a_bufs) == len(b_bufs) == len(c_bufs) == n):
        raise ValueError("plans and buffers differ in number")
    abc = [p._abc(a, b, c) for p, a, b, c in zip(plans, a_bufs, b_bufs, c_bufs)]
    arr = lambda xs: (_vp * n)(*xs)  # noqa: E731
    return (n, arr([p.handle for p in plans]), arr([x[0] for x in abc]), arr([x[1] for x in abc]),
            arr([x[2] for x in abc]))


def batch_multiply_device(plans, a_bufs, b_bufs, c_bufs, stream=None) -> None:
    _check(lib.pm_batch_multiply_device(*_batch_args(plans, a_bufs, b_bufs, c_bufs), _sv(stream)))


def batch_multiply_host(plans, a_list, b_list, depth=0, out=None):
    """Products of numpy operands through the plans' host pipelines, up to
    `depth` at once (pm_batch_multiply_host); returns the answers (or fills
    the arrays in `out`)."""
    n = len(plans)
    if not (len(a_list) == len(b_list) == n):
        raise ValueError("plans and operands differ in number")
    a_list = [p._operand(a) for p, a in zip(plans, a_list)]
    b_list = [p._operand(b) for p, b in zip(plans, b_list)]
    if out is None:
        out = [np.empty((p.N, p.L), dtype=np.uint32) for p in plans]
    elif len(out) != n or any(not (isinstance(c, np.ndarray) and c.dtype == np.uint32 and c.shape == (p.N, p.L)
                                   and c.flags["C_CONTIGUOUS"]) for p, c in zip(plans, out)):
        raise ValueError("out must hold one C-contiguous uint32 (N, L) array per product")
    arr = lambda xs: (_vp * n)(*xs)  # noqa: E731
    addr = lambda xs: arr([x.ctypes.data for x in xs])  # noqa: E731
    _check(lib.pm_batch_multiply_host(n, arr([p.handle for p in plans]), addr(a_list), addr(b_list), addr(out),
                                      int(depth)))
    return out


class Batch:
    """A batch of products captured once as a hipGraph; launch() replays it.

    Without `shape`, creation times candidate shapes and keeps the fastest
    (pm_batch_create); with shape = (runs, streams, ntt_form, twiddles) it
    captures that one (pm_batch_create_shaped)."""

    def __init__(self, plans, a_bufs, b_bufs, c_bufs, shape=None):
        h = _vp()
        args = _batch_args(plans, a_bufs, b_bufs, c_bufs)
        if shape is None:
            _check(lib.pm_batch_create(*args, ctypes.byref(h)))
        else:
            sh = BatchShape(*[int(x) for x in shape])
            _check(lib.pm_batch_create_shaped(*args, ctypes.byref(sh), ctypes.byref(h)))
        self._h = h.value
        self._keep = (list(plans), list(a_bufs), list(b_bufs), list(c_bufs))

    @property
    def shape(self):
        """(runs, streams, ntt_form, twiddles) as captured."""
        sh = BatchShape()
        _check(lib.pm_batch_get_shape(_vp(self._h), ctypes.byref(sh), None))
        return (sh.runs, sh.streams, sh.ntt_form, sh.twiddles)

    @property
    def tuned_ms(self) -> float:
        """Median replay (ms) of the shape the timed creation chose, else 0."""
        ms = ctypes.c_double()
        _check(lib.pm_batch_get_shape(_vp(self._h), None, ctypes.byref(ms)))
        return ms.value

    def launch(self, stream=None) -> None:
        _check(lib.pm_batch_launch(_vp(self._h), _sv(stream)))

    @property
    def phased(self) -> bool:
        """True if captured as phase-batched launches (one kernel per phase for all products)."""
        return bool(lib.pm_batch_phased(_vp(self._h)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.pm_batch_destroy(_vp(self._h))
            self._h = None

    def __del__(self):
        self.close()
