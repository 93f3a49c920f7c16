"""Build recipe for libfherma_polymult.so (gfx950): every HIP and C++ source in
csrc/ compiled by hipcc, objects in _build/, the library beside this file.

    python fherma-polymult_amd/build.py
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfherma_polymult.so")
SOLUTION = os.path.join(HERE, "solution")   # point-directory driver (harness/solution_main.cpp)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are switched off for all device code: the kernels
# use none on purpose, and the compiler must not form them either. The matrix
# cores accumulate in ordinary VGPRs (gfx950 allows it): the MFMA kernels'
# epilogues then read their sums without accumulator-register copies.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-Wall", "-Wno-unused-function", "-I" + os.path.join(HERE, "..", "include")]
# The NTT and matrix-core kernels raise each wave's priority (s_setprio) until
# its last global load is issued, so that a wave starting a block sends its loads
# ahead of the other waves' arithmetic: the inverse column pass alone 17.1 -> 16.6
# us, the 16-product batch 0.2-0.3 % faster (ABBA on one box).
# The NTT kernels are scheduled for instruction-level parallelism rather than
# occupancy: the butterfly code issues faster from fewer waves with more
# independent work interleaved (the fused row pass at 124 VGPRs and four waves
# per SIMD against 69 and seven: 56.9 -> 53.0 us alone, the batch +0.8-1.9 %).
FILE_FLAGS = {"pm_ntt.hip": ["-mllvm", "-amdgpu-set-wave-priority", "-mllvm", "-amdgpu-sched-strategy=max-ilp"],
              "pm_mfma.hip": ["-mllvm", "-amdgpu-set-wave-priority"]}


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    if not src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), out))
    return obj, out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    for _, out in results:
        if verbose and out:
            print(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    # The driver is host-only C++ against the FHERMA boundary: g++ with the same
    # libstdc++ ABI as the library, found next to it at run time.
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "harness", "solution_main.cpp"), "-o", SOLUTION + ".tmp",
           "-L" + HERE, "-lfherma_polymult", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(SOLUTION + ".tmp", SOLUTION)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
