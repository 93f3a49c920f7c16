"""CPU checks of the product library: it loads, exports every entry point its
headers declare (the C ABI and the FHERMA C++ symbols), and its host-side plan
logic agrees with the oracle. No kernel is launched here."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
import polymult

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", polymult.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def declared_c_functions():
    text = open(os.path.join(ROOT, "include", "fherma_polymult.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text))


def test_every_declared_c_function_is_exported():
    syms = exported_symbols()
    declared = declared_c_functions()
    assert len(declared) >= 25
    missing = sorted(d for d in declared if d not in syms)
    assert not missing, missing
    for name in declared:
        getattr(polymult.lib, name)   # ctypes resolves it


def test_fherma_cpp_entry_points_are_exported():
    # The reference's boundary has C++ linkage (fherma.h:46-48): these are the
    # symbols a harness built from its main.cpp binds.
    syms = exported_symbols()
    for mangled in ("_Z11fherma_initRKN6fherma5PointE", "_Z10fherma_runPvRKN6fherma6InputsE", "_Z11fherma_freePv"):
        assert mangled in syms, mangled


@pytest.mark.parametrize("N,W", [(2, 3), (64, 2048), (4096, 256), (32768, 868), (65536, 1747), (65536, 2048),
                                 (1 << 17, 512), (1 << 22, 32)])
def test_host_prime_rule_matches_oracle(N, W):
    assert np.array_equal(polymult.choose_primes(N, W), oracle.choose_primes(N, W))


def test_prime_rule_refusals():
    """The reference's messages (rns_plan.h:120-155): too few primes of the form
    for the point, and an N that is not a power of two."""
    with pytest.raises(polymult.PolymulError) as e:
        polymult.choose_primes(1 << 22, 2048)
    assert e.value.code == polymult.PM_EINVAL and "not enough primes p < 2^31" in str(e.value)
    with pytest.raises(polymult.PolymulError) as e:
        polymult.choose_primes(3000, 256)
    assert e.value.code == polymult.PM_EINVAL


def test_invalid_points_are_refused_before_any_device_work():
    q = np.array([5], dtype=np.uint32)
    for N in (0, 1, 3, 1 << 23):
        with pytest.raises(polymult.PolymulError) as e:
            polymult.Plan(N, 3, q)
        assert e.value.code == polymult.PM_EINVAL
    with pytest.raises(polymult.PolymulError) as e:
        polymult.Plan(64, 2049, np.ones(65, dtype=np.uint32))
    assert e.value.code == polymult.PM_EINVAL
    with pytest.raises(polymult.PolymulError) as e:
        polymult.Plan(64, 64, np.array([7, 0], dtype=np.uint32))
    assert e.value.code == polymult.PM_EINVAL
    assert "fewer limbs" in str(e.value)


def test_not_enough_primes_is_refused():
    # N = 2^22 leaves primes = 1 mod 2^23 below 2^31: far fewer than 2048 bits need.
    with pytest.raises(polymult.PolymulError) as e:
        polymult.choose_primes(1 << 22, 2048)
    assert e.value.code == polymult.PM_EINVAL
    with pytest.raises(ValueError):
        oracle.choose_primes(1 << 22, 2048)


def test_driver_refuses_out_of_range_points(tmp_path):
    """N = 2^23 and W = 2049 (L = 65) are refused by fherma_init before any
    device work: non-zero exit, no answers (reference fuzz.py:149-160)."""
    import json
    solution = os.path.join(ROOT, "fherma-polymult_amd", "solution")
    for N, W, L in ((1 << 23, 32, 1), (64, 2049, 65)):
        d = tmp_path / ("p%d_%d" % (N, W))
        (d / "point").mkdir(parents=True)
        np.ones(L, dtype="<u4").tofile(str(d / "point" / "q.bin"))
        (d / "manifest.json").write_text(json.dumps({"N": N, "W": W, "L": L, "cases": 1}))
        r = subprocess.run([solution, str(d)], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0
        assert not (d / "out" / "000000").exists()


def test_driver_refuses_unreadable_point_directories(tmp_path):
    """A missing manifest, a missing or malformed value, or a q.bin of the wrong
    size ends the driver with status 2 and a message, before fherma_init."""
    import json
    solution = os.path.join(ROOT, "fherma-polymult_amd", "solution")
    good = {"N": 64, "W": 32, "L": 1, "cases": 1}
    bad = [None, {"N": 64, "W": 32, "L": 1}, dict(good, N=-64), dict(good, cases=1.5), "{\"N\" 64}"]
    for i, manifest in enumerate(bad + ["short-q"]):
        d = tmp_path / ("d%d" % i)
        (d / "point").mkdir(parents=True)
        np.ones(1, dtype="<u4")[:0 if manifest == "short-q" else 1].tofile(str(d / "point" / "q.bin"))
        if manifest == "short-q":
            manifest = good
        if isinstance(manifest, dict):
            (d / "manifest.json").write_text(json.dumps(manifest))
        elif manifest is not None:
            (d / "manifest.json").write_text(manifest)
        r = subprocess.run([solution, str(d)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and r.stderr.startswith("solution: "), (manifest, r.returncode, r.stderr)
        assert not (d / "out" / "results.json").exists()


def test_invalid_twiddle_policy_is_refused_before_device_work():
    """FHERMA_TWIDDLES other than table, gen or auto fails plan creation with
    PM_EINVAL, as the reference's fherma_init does (rns_plan.h:273-280)."""
    code = ("import sys; sys.path.insert(0, %r); import numpy as np, polymult\n"
            "try:\n"
            "    polymult.Plan(64, 64, np.array([7, 1], dtype=np.uint32))\n"
            "except polymult.PolymulError as e:\n"
            "    print(e.code, str(e))\n" % os.path.join(ROOT, "fherma-polymult_amd"))
    env = dict(os.environ, FHERMA_TWIDDLES="fast")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.stdout.startswith("%d " % polymult.PM_EINVAL), out.stdout + out.stderr
    assert "FHERMA_TWIDDLES" in out.stdout


def test_new_entry_points_refuse_bad_arguments_on_the_host():
    """The batch, host-batch, shape, sum and form entry points check their
    arguments before any device work (no GPU is needed for these)."""
    import ctypes
    lib = polymult.lib
    vp = ctypes.c_void_p
    null_arr = ctypes.POINTER(vp)()
    assert lib.pm_batch_multiply_host(0, null_arr, null_arr, null_arr, null_arr, 0) == polymult.PM_OK
    assert lib.pm_batch_multiply_host(2, null_arr, null_arr, null_arr, null_arr, 0) == polymult.PM_EINVAL
    two = (vp * 2)(None, None)
    assert lib.pm_batch_multiply_host(2, two, two, two, two, 1) == polymult.PM_EINVAL
    assert "null" in lib.pm_last_error().decode()
    out = vp()
    shape = polymult.BatchShape(1, 1, -1, -1)
    assert lib.pm_batch_create_shaped(0, two, two, two, two, ctypes.byref(shape), ctypes.byref(out)) == polymult.PM_EINVAL
    assert lib.pm_batch_create(1, two, two, two, two, None) == polymult.PM_EINVAL
    assert lib.pm_batch_get_shape(None, ctypes.byref(shape), None) == polymult.PM_EINVAL
    s = ctypes.c_uint64()
    assert lib.pm_device_sum(None, 16, ctypes.byref(s)) == polymult.PM_EINVAL
    assert lib.pm_plan_set_ntt_form(None, 0) == polymult.PM_EINVAL
    assert lib.pm_plan_get_ntt_form(None) == -1
    ms = ctypes.c_double()
    assert lib.pm_plan_tune_kernels(None, ctypes.byref(ms)) == polymult.PM_EINVAL
