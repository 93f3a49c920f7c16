"""pm_multiply_device_noalloc: the product with its residue workspace in the
library's static device array (no allocation in the call), against
pm_multiply_device on the same operands, bit for bit, and against the GMP
digests; two calls in a row, the second on a smaller point, then the first
point again; and a point whose workspace exceeds the static array is refused
before any launch."""
import sys

import numpy as np
import pytest

import oracle
import pointgen

polymult = pytest.importorskip("polymult")

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def run_both(case, form):
    q, a, b = pointgen.operands(case)
    with polymult.Plan(case["N"], case["W"], q, form=form) as plan:
        da, db = polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b)
        c_dev, c_na = polymult.DeviceBuffer(a.nbytes), polymult.DeviceBuffer(a.nbytes)
        c_na.upload(np.full(a.size, 0xA5A5A5A5, dtype=np.uint32))
        plan.multiply_device(da, db, c_dev)
        plan.multiply_device_noalloc(da, db, c_na)
        polymult.synchronize()
        want = c_dev.download((plan.N, plan.L))
        got = c_na.download((plan.N, plan.L))
        for x in (da, db, c_dev, c_na):
            x.close()
    assert np.array_equal(got, want), case["name"]
    assert oracle.digest(got) == case["c_sha"], case["name"]


@pytest.mark.parametrize("form", [1, 0], ids=["mfma", "plain"])
def test_noalloc_matches_device_entry_point(golden, form):
    big, small = golden["config-n65536-w2048-00"], golden["config-n4096-w256-s1"]
    for case in (big, small, big):
        say("noalloc", case["name"], "form", form)
        run_both(case, form)


def test_noalloc_refuses_a_workspace_beyond_the_static_array():
    # a 2048-bit q at N = 131072 needs 2 x 134 x 131072 words
    q2, _, _ = oracle.gen_point(17, 131072, 2048)
    with polymult.Plan(131072, 2048, q2) as plan:
        assert 2 * plan.K * plan.N > 72 << 18
        d = polymult.DeviceBuffer(4 * plan.N * plan.L)
        with pytest.raises(polymult.PolymulError) as e:
            plan.multiply_device_noalloc(d, d, d)
        assert e.value.code == polymult.PM_EINVAL and "workspace" in str(e.value)
        d.close()
