"""Batches of uneven products on the GPU: batched launches over products of
one N with different K and L, every captured shape, the timed shape choice,
and the whole 64-product uneven workload, each product against its GMP digest.
"""
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


class Products:
    """Plans and device buffers for golden cases; closed together."""

    def __init__(self, golden, names, form=None):
        self.cases = [golden[n] for n in names]
        self.plans, self.da, self.db, self.dc = [], [], [], []
        for case in self.cases:
            q, a, b = pointgen.operands(case)
            f = polymult.FORM_MFMA if form is None else form
            self.plans.append(polymult.Plan(case["N"], case["W"], q, form=f))
            self.da.append(polymult.DeviceBuffer.from_array(a))
            self.db.append(polymult.DeviceBuffer.from_array(b))
            self.dc.append(polymult.DeviceBuffer(a.nbytes))

    def clear(self):
        for c in self.dc:
            c.upload(np.zeros(c.nbytes // 4, dtype=np.uint32))

    def check(self, tag):
        bad = [case["name"] for case, c in zip(self.cases, self.dc)
               if oracle.digest(c.download((case["N"], case["L"]))) != case["c_sha"]]
        assert not bad, (tag, bad)

    def close(self):
        for x in self.da + self.db + self.dc + self.plans:
            x.close()


def uneven(golden):
    return sorted(n for n in golden if n.startswith("config-uneven-"))


def test_whole_uneven_workload_tuned_graph_matches_gmp(golden):
    """The 64 products of uneven N (4096 .. 65536) and q width (128 .. 2019 bits),
    captured with the timed shape choice; three replays, each checked."""
    names = uneven(golden)
    assert len(names) == 64
    ps = Products(golden, names)
    say("uneven workload: tuning the batch of", len(names))
    batch = polymult.Batch(ps.plans, ps.da, ps.db, ps.dc)
    runs, streams, ntt, tw = batch.shape
    say("chosen shape", batch.shape, "median ms", batch.tuned_ms)
    assert batch.tuned_ms > 0
    assert runs >= 0 and 1 <= streams <= 64
    for rep in range(3):
        ps.clear()
        batch.launch()
        polymult.synchronize()
        ps.check(("replay", rep))
    batch.close()
    ps.close()


SHAPES = [(0, 4, -1, -1), (0, 1, -1, -1), (1, 1, -1, -1), (5, 5, -1, -1), (7, 2, 0, 1), (16, 16, 0, 0),
          (16, 4, 1, 1), (40, 16, 1, 0)]


@pytest.mark.parametrize("shape", SHAPES, ids=["r%d-s%d-n%d-t%d" % s for s in SHAPES])
def test_every_shape_matches_gmp(golden, shape):
    """28 uneven products (every N, K from 10 to 132) under each captured
    shape: chains only, one run, runs per N, split runs, both column-pass forms
    and both twiddle sources."""
    names = uneven(golden)[::2][:28]
    ps = Products(golden, names)
    say("shape", shape, "products", len(names))
    batch = polymult.Batch(ps.plans, ps.da, ps.db, ps.dc, shape=shape)
    assert batch.tuned_ms == 0
    got = batch.shape
    assert got[0] == shape[0] and got[2:] == shape[2:] and 1 <= got[1] <= shape[1]
    assert batch.phased == (shape[0] > 0)
    for rep in range(2):
        ps.clear()
        batch.launch()
        polymult.synchronize()
        ps.check((shape, rep))
    batch.close()
    ps.close()


def test_batched_launches_on_a_stream_match_gmp(golden):
    """pm_batch_multiply_device on an explicit stream: the uneven products go
    out as batched launches per N (decompose and CRT per K-step count)."""
    names = uneven(golden)[1::2]
    ps = Products(golden, names)
    stream = polymult.Stream()
    say("batched on a stream", len(names))
    polymult.batch_multiply_device(ps.plans, ps.da, ps.db, ps.dc, stream=stream)
    stream.synchronize()
    ps.check("stream")
    ps.close()


def test_mixed_batch_alone_products_match_gmp(golden):
    """Products that cannot share launches (the plain form; N = 2^17, whose
    geometry has no register kernels) beside ones that can, in one graph."""
    names = uneven(golden)[:10] + ["matrix-n131072-w512-prime", "config-n4096-w256-s1"]
    ps = Products(golden, names)
    extra = golden["config-n32768-w868-s1"]
    q, a, b = pointgen.operands(extra)
    ps.cases.append(extra)
    ps.plans.append(polymult.Plan(extra["N"], extra["W"], q, form=polymult.FORM_PLAIN))
    ps.da.append(polymult.DeviceBuffer.from_array(a))
    ps.db.append(polymult.DeviceBuffer.from_array(b))
    ps.dc.append(polymult.DeviceBuffer(a.nbytes))
    for shape in ((5, 4, -1, -1), (0, 4, -1, -1)):
        say("mixed batch", shape)
        batch = polymult.Batch(ps.plans, ps.da, ps.db, ps.dc, shape=shape)
        ps.clear()
        batch.launch()
        polymult.synchronize()
        ps.check(shape)
        batch.close()
    stream = polymult.Stream()
    ps.clear()
    polymult.batch_multiply_device(ps.plans, ps.da, ps.db, ps.dc, stream=stream)
    stream.synchronize()
    ps.check("stream")
    ps.close()


def test_tuner_environment_overrides(golden, monkeypatch):
    names = uneven(golden)[:12]
    ps = Products(golden, names)
    monkeypatch.setenv("PM_BATCH_TUNE", "0")
    batch = polymult.Batch(ps.plans, ps.da, ps.db, ps.dc)
    assert batch.tuned_ms == 0 and batch.shape[0] == 16 and batch.shape[2:] == (-1, -1)
    batch.close()
    monkeypatch.delenv("PM_BATCH_TUNE")
    monkeypatch.setenv("PM_BATCH_RUNS", "3")
    monkeypatch.setenv("PM_BATCH_STREAMS", "2")
    monkeypatch.setenv("PM_NTT_FORM", "valu")
    monkeypatch.setenv("FHERMA_TWIDDLES", "gen")
    say("tuner with runs and streams fixed")
    batch = polymult.Batch(ps.plans, ps.da, ps.db, ps.dc)
    assert batch.tuned_ms > 0 and batch.shape == (3, 2, -1, -1)
    ps.clear()
    batch.launch()
    polymult.synchronize()
    ps.check("overrides")
    batch.close()
    monkeypatch.setenv("PM_BATCH_RUNS", "many")
    with pytest.raises(polymult.PolymulError) as e:
        polymult.Batch(ps.plans, ps.da, ps.db, ps.dc)
    assert e.value.code == polymult.PM_EINVAL
    ps.close()


def test_full_twiddle_tables_on_demand(golden):
    """A plan whose rule picks generated row twiddles holds no full tables;
    switching to tables builds them, back to generated releases them, and
    every setting gives the GMP product."""
    case = golden["config-n65536-w2048-01"]
    q, a, b = pointgen.operands(case)
    with polymult.Plan(case["N"], case["W"], q) as plan:
        assert plan.twiddles == polymult.TWIDDLES_GEN
        for source in (polymult.TWIDDLES_TABLE, polymult.TWIDDLES_GEN, polymult.TWIDDLES_TABLE):
            plan.set_twiddles(source)
            say("twiddles", source)
            assert oracle.digest(plan.multiply(a, b)) == case["c_sha"], source


def test_more_than_32_uneven_products_of_one_n_match_single_products():
    """40 products at N = 4096 with q widths from 100 to 2048 bits (K from 9 to
    127, L from 4 to 64): one run is cut into launches of 32 and 8, each with
    its own plan of most primes and its own K-step buckets; the results equal
    each product computed alone, through the graph and on a stream."""
    rng = np.random.default_rng(4096)
    N = 4096
    widths = [int(w) for w in rng.integers(100, 2049, 40)]
    pts = [oracle.gen_point(9000 + i, N, w) for i, w in enumerate(widths)]
    plans = [polymult.Plan(N, w, q) for (q, _, _), w in zip(pts, widths)]
    assert len({p.K for p in plans}) > 10 and len({p.L for p in plans}) > 10
    da = [polymult.DeviceBuffer.from_array(a) for _, a, _ in pts]
    db = [polymult.DeviceBuffer.from_array(b) for _, _, b in pts]
    ref = [polymult.DeviceBuffer(a.nbytes) for _, a, _ in pts]
    out = [polymult.DeviceBuffer(a.nbytes) for _, a, _ in pts]
    for p, a, b, c in zip(plans, da, db, ref):
        p.multiply_device(a, b, c)
    polymult.synchronize()
    want = [c.download((N, p.L)) for c, p in zip(ref, plans)]
    # the single products themselves against the oracle, for three of them
    for i in (0, 17, 39):
        q, a, b = pts[i]
        assert np.array_equal(want[i], oracle.Plan(N, q).multiply(a, b)), i
    say("40 uneven products at one N: one run")
    batch = polymult.Batch(plans, da, db, out, shape=(1, 1, -1, -1))
    assert batch.phased
    batch.launch()
    polymult.synchronize()
    for i, (c, p) in enumerate(zip(out, plans)):
        assert np.array_equal(c.download((N, p.L)), want[i]), ("graph", i)
    batch.close()
    stream = polymult.Stream()
    for c in out:
        c.upload(np.zeros(c.nbytes // 4, dtype=np.uint32))
    polymult.batch_multiply_device(plans, da, db, out, stream=stream)
    stream.synchronize()
    for i, (c, p) in enumerate(zip(out, plans)):
        assert np.array_equal(c.download((N, p.L)), want[i]), ("stream", i)
    for x in da + db + ref + out + plans:
        x.close()
