"""The library's host plan (fherma-polymult_amd/csrc/pm_plan.cpp) built for the
CPU with g++ and a C shim (tests/host_plan_shim.cpp): the constants a plan
would upload, without a device. Used by tests/test_host_plan.py and by the CRT
kernel model's tests, which need the exact FP64 ratios the kernels read."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fherma-polymult_amd", "csrc")
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(outdir):
    """Compiles pm_plan.cpp and the shim into outdir; returns the loaded library."""
    lib = os.path.join(outdir, "libhostplan.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I" + CSRC,
                    os.path.join(CSRC, "pm_plan.cpp"), os.path.join(ROOT, "tests", "host_plan_shim.cpp"),
                    "-o", lib], check=True)
    L = ctypes.CDLL(lib)
    L.hp_plan.restype = ctypes.c_int
    L.hp_plan.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, ctypes.c_int] + [_u32p] * 9 + \
                         [_f64p, _f64p, ctypes.c_char_p, ctypes.c_int]
    return L


def plan_of(hp, N, W, q):
    """Every constant of the host plan for (N, W, q) as numpy arrays; ValueError
    with the library's message for a refused point."""
    L = q.shape[0]
    kmax = 200
    bufs = {k: np.zeros(n, dtype=np.uint32) for k, n in
            (("p", kmax), ("pneg", kmax), ("psi", kmax), ("ipsi", kmax), ("scale", kmax),
             ("pow2", kmax * L), ("pow2b", kmax * L), ("M", kmax * L), ("negvP", (kmax + 1) * L))}
    mq, nq = np.zeros(kmax), np.zeros(kmax + 1)
    err = ctypes.create_string_buffer(256)
    u = lambda a: a.ctypes.data_as(_u32p)  # noqa: E731
    K = hp.hp_plan(N, W, u(q), L, u(bufs["p"]), u(bufs["pneg"]), u(bufs["psi"]), u(bufs["ipsi"]),
                   u(bufs["scale"]), u(bufs["pow2"]), u(bufs["pow2b"]), u(bufs["M"]), u(bufs["negvP"]),
                   mq.ctypes.data_as(_f64p), nq.ctypes.data_as(_f64p), err, 256)
    if K < 0:
        raise ValueError(err.value.decode())
    out = {k: v[:K] for k, v in bufs.items() if k in ("p", "pneg", "psi", "ipsi", "scale")}
    for k in ("pow2", "pow2b", "M"):
        out[k] = bufs[k][:K * L].reshape(K, L)
    out["negvP"] = bufs["negvP"][:(K + 1) * L].reshape(K + 1, L)
    out["mq"], out["nq"], out["K"] = mq[:K], nq[:K + 1], K
    return out


def limbs(x, L):
    return np.array([(x >> (32 * i)) & 0xFFFFFFFF for i in range(L)], dtype=np.uint32)


def value(row):
    return sum(int(w) << (32 * i) for i, w in enumerate(row))
