#!/usr/bin/env python3
"""Throughput of the exact negacyclic product on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form mfma|plain] [--no-cpu]

Workload: a batch of 16 independent products c = a b mod (X^N + 1, q) with
N = 65536 and a seeded random odd 2048-bit q per product (64 u32 limbs, K = 134
primes each), a and b uniform in [0, q). One step = the whole batch, device
resident (inputs already in HBM when the timed region starts). Prints one JSON
line: products/s of the matrix-core form (value) beside the plain form, the
roofline of the dominant kernel from HIP-event timings, and the CPU baseline
(the reference's solve_cpu.cpp, or the oracle's port of it where that library
is absent) on a bounded sample. Every product of the batch is checked against
its GMP digest in tests/golden/golden.json.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fherma-polymult_amd"))

N, W, BATCH, SEED0 = 65536, 2048, 16, 5000
HBM_PEAK_GBS = 8000.0
G = 0x9E3779B97F4A7C15
H = 0xD1B54A32D192ED03
M64 = (1 << 64) - 1


# --- seeded inputs (counter-based SplitMix64; the fixtures use the same stream)
def _mix_int(z):
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _words(key, idx):
    with np.errstate(over="ignore"):
        z = np.uint64(key) + (idx + np.uint64(1)) * np.uint64(G)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(32)).astype(np.uint32)


def make_point(seed, n, w):
    L = (w + 31) // 32
    mask = np.uint32((1 << (w % 32)) - 1 if w % 32 else 0xFFFFFFFF)
    q = _words(_mix_int(seed * G), np.arange(L, dtype=np.uint64))
    q[L - 1] &= mask
    q[L - 1] |= np.uint32(1 << ((w - 1) % 32))
    q[0] |= np.uint32(1)
    return make_operands(seed, n, w, q)


def make_operands(seed, n, w, q):
    """(q, a, b): a and b uniform in [0, q) from the seed's streams 1 and 2."""
    L = (w + 31) // 32
    mask = np.uint32((1 << (w % 32)) - 1 if w % 32 else 0xFFFFFFFF)

    def operand(stream):
        key = _mix_int(seed * G + stream * H)
        out = np.zeros((n, L), dtype=np.uint32)
        pending = np.arange(n, dtype=np.uint64)
        lanes = np.arange(L, dtype=np.uint64)
        for t in range(64):
            if pending.size == 0:
                break
            cand = _words(key, ((pending * np.uint64(64) + np.uint64(t)) * np.uint64(L))[:, None] + lanes[None, :])
            cand[:, L - 1] &= mask
            lt = np.zeros(pending.size, dtype=bool)
            eq = np.ones(pending.size, dtype=bool)
            for l in range(L - 1, -1, -1):
                lt |= eq & (cand[:, l] < q[l])
                eq &= cand[:, l] == q[l]
            out[pending[lt].astype(np.int64)] = cand[lt]
            pending = pending[~lt]
        return out

    return q, operand(1), operand(2)


def digest(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u4").tobytes()).hexdigest()


def golden_digests():
    path = os.path.join(ROOT, "tests", "golden", "golden.json")
    with open(path) as f:
        cases = json.load(f)["cases"]
    return {c["seed"]: c["c_sha"] for c in cases if c["name"].startswith("config-n65536-w2048-")}


# --- bytes each kernel must read and write per launch (K x N residues, N x L limbs)
def phase_bytes(K, L):
    kn, nl = 4 * K * N, 4 * N * L
    return {"decompose_a": nl + kn, "decompose_b": nl + kn, "ntt_cols_fwd_a": 2 * kn, "ntt_cols_fwd_b": 2 * kn,
            "ntt_rows_fused": 3 * kn, "ntt_cols_inv": 2 * kn, "crt": kn + nl}


def phase_kernels(plan):
    """Kernel names (as rocprofv3 reports them, spaces removed) of each phase of
    the N = 65536 one-product schedule: the register NTT kernels take the twiddle
    source and the butterfly form (1: forward residues in [0, 2p)) as template
    arguments, the CRT its K-steps (one group per four primes plus the quotient
    group, 16 groups per K-step)."""
    lazy = "1"
    gen = "true" if plan.twiddles == 1 else "false"
    ks = ((plan.K + 3) // 4 + 1 + 3) // 4
    col = "k_ntt_col_reg<8,%s,5," + lazy + ">"
    if plan.fused_cols == 1:
        # the first four column stages inside decompose, the next four in the row pass
        return {"decompose_a": "k_decompose_cols_batch<16,4,1>@1", "decompose_b": "k_decompose_cols_batch<16,4,1>@1",
                "ntt_rows_fused": "k_ntt_rows_fused_mid_batch<%s,%s>" % (gen, lazy),
                "ntt_cols_inv": col % "false", "crt": "k_crt_mfma<%d>" % ks}
    return {"decompose_a": "k_decompose_mfma<4,true>", "decompose_b": "k_decompose_mfma<4,true>",
            "ntt_cols_fwd_a": col % "true", "ntt_cols_fwd_b": col % "true",
            "ntt_rows_fused": "k_ntt_rows_fused_reg<8,%s,%s>" % (gen, lazy), "ntt_cols_inv": col % "false",
            "crt": "k_crt_mfma<%d>" % ks}


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed counter table
    (profiles/hbm_traffic_n65536_w2048.txt: kernel, scaled FETCH_SIZE + WRITE_SIZE
    bytes per dispatch), or None when the table or the kernel is absent."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_n65536_w2048.txt")
    if not os.path.exists(path):
        return None
    want = kernel
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and parts[0] == want and parts[1].isdigit():
                return int(parts[1])
    return None


# Cycles per wave64 instruction per SIMD on gfx950, from the issue-rate probe
# (pm_probe_issue; profiles/issue_rates_gfx950.md): plain integer ops, 32-bit
# multiplies (low and high half), 32 x 32 -> 64-bit multiply-adds; and the
# MI355X peak engine clock.
ISSUE_CYCLES = {"other": 2.92, "mul_lo": 5.04, "mul_hi": 4.64, "mad64": 5.9}
CLOCK_GHZ = 2.4
# VALU instructions per wave of the NTT kernels by class, counted in the gfx950
# assembly of this build (straight-line kernels: the static count is the
# per-wave count; it matches SQ_INSTS_VALU per wave).
VALU_MIX = {
    "k_ntt_rows_fused_reg<8,true,1>": {"other": 1584, "mul_hi": 448, "mul_lo": 321, "mad64": 1},
    # with the four column stages of distance 256 .. 2048 on A and B before the rows
    "k_ntt_rows_fused_mid_batch<true,1>": {"other": 2055, "mul_hi": 512, "mul_lo": 448},
    "k_ntt_col_reg<8,true,5,1>": {"other": 465, "mul_hi": 64, "mul_lo": 128},
    "k_ntt_col_reg<8,false,5,1>": {"other": 465, "mul_hi": 128, "mul_lo": 64},
}


def valu_issue(kernel, kernel_ms, cus):
    """The VALU issue time of one launch of a kernel: its instructions per wave
    by class (VALU_MIX) at the probe's rates, times its waves per launch from the
    committed SQ counter table (profiles/sq_counters_n65536_w2048.txt), over
    the SIMDs; None when either is absent."""
    path = os.path.join(ROOT, "profiles", "sq_counters_n65536_w2048.txt")
    mix = VALU_MIX.get(kernel)
    if mix is None or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == kernel:
                waves, per_wave = int(parts[1]), float(parts[2])
                simds = 4 * cus
                cycles = sum(n * ISSUE_CYCLES[k] for k, n in mix.items())
                ms = cycles * waves / simds / (CLOCK_GHZ * 1e9) * 1e3
                return {"instructions_per_wave": per_wave, "mix": mix, "cycles_per_wave": round(cycles, 1),
                        "cycles_per_instruction": round(cycles / sum(mix.values()), 3), "waves": waves,
                        "clock_ghz": CLOCK_GHZ, "simds": simds, "issue_ms": round(ms, 5),
                        "frac": round(ms / kernel_ms, 4),
                        "note": "issue time of the kernel's VALU instructions at the probe's rates (independent "
                                "chains) over its measured time; copies of the row pass without loads, stores or "
                                "LDS exchanges kept most of the rest, so it is the issue of the butterflies' "
                                "dependent chains, not memory"}
    return None


def host_cpus():
    """(CPUs this process may run on, CPUs the host has)."""
    try:
        mine = len(os.sched_getaffinity(0))
    except AttributeError:
        mine = os.cpu_count() or 1
    return mine, os.cpu_count() or 1


def cpu_baseline(points, budget_s=12.0):
    """The reference's CPU path (oracle/_ref) on the host cores this process may
    use (OMP_NUM_THREADS where the environment sets a share), else the oracle port."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402  (the checker's CPU port, used only for this leg)
    mine, total = host_cpus()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or mine
    oracle.set_threads(threads)
    q, a, b = points[0]
    if oracle.ref_available():
        kind, solver = "reference", oracle.RefSolver(N, W, q)
        run = solver.run
    else:
        kind, plan = "port", oracle.Plan(N, q)
        run = plan.multiply
    run(a, b)   # warm
    done, t0 = 0, time.perf_counter()
    while True:
        run(a, b)
        done += 1
        el = time.perf_counter() - t0
        if el > budget_s or done >= 64:
            break
    return {"value": done / el, "unit": "polymuls/s", "cores": threads, "kind": kind,
            "host_cpus": total, "cpus_available": mine,
            "sample": "%d products at N=65536, W=2048 (K=134), %.1f s, OMP threads=%d of %d CPUs available "
                      "to this process (%d on the host)" % (done, el, threads, mine, total)}


def batch_rates(args, plans, da, db, dc, gold, log):
    """Products/s of the matrix-core batch (one hipGraph) and of the plain form."""
    import polymult  # noqa: E402
    L = plans[0].L
    t0 = time.perf_counter()
    graph = polymult.Batch(plans, da, db, dc)
    shape = shape_info(graph, time.perf_counter() - t0)
    log("batch shape (timed choice): %s" % shape)

    def step():
        graph.launch()

    def timed(steps, warmup):
        for _ in range(warmup):
            step()
        polymult.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        polymult.synchronize()
        return (time.perf_counter() - t0) / steps

    # correctness of the batch against the GMP digests
    step()
    polymult.synchronize()
    mismatches = sum(digest(dc[i].download((N, L))) != gold.get(SEED0 + i) for i in range(BATCH))
    log("mfma batch: %d mismatches vs GMP" % mismatches)

    sec = timed(args.steps, args.warmup)
    value = BATCH / sec
    log("mfma: %.3f ms per step, %.1f products/s" % (sec * 1e3, value))

    # The plain HIP form: VALU decompose and CRT, products launched one by one.
    plain_value = None
    if not args.no_plain:
        for p in plans:
            p.set_form(polymult.FORM_PLAIN)

        def step():  # noqa: F811
            polymult.batch_multiply_device(plans, da, db, dc)

        step()
        polymult.synchronize()
        mismatches += sum(digest(dc[i].download((N, L))) != gold.get(SEED0 + i) for i in range(BATCH))
        plain_sec = timed(max(3, args.steps // 4), 1)
        plain_value = BATCH / plain_sec
        log("plain: %.3f ms per step, %.1f products/s" % (plain_sec * 1e3, plain_value))
        for p in plans:
            p.set_form(polymult.FORM_MFMA)

    return value, sec, plain_value, mismatches, shape


def shape_info(graph, create_s):
    """The captured shape of a timed batch, for the bench line."""
    runs, streams, ntt, tw = graph.shape
    names = {-1: "plans' own", 0: "valu", 1: "mfma"}
    return {"runs": runs, "streams": streams, "ntt_form": names[ntt],
            "twiddles": {-1: "plans' own", 0: "table", 1: "generated"}[tw],
            "tuned_replay_ms": round(graph.tuned_ms, 4), "create_s": round(create_s, 3)}


def short_kernel(name):
    """A rocprofv3 kernel name without namespaces, parameters and spaces."""
    return name[name.find("k_"):].split("(")[0].replace(" ", "")


def trace_stats(path, log):
    """Steady-state durations per kernel from a rocprofv3 kernel trace (CSV):
    the median and the 10 %-trimmed mean drop the first launches' warm-up."""
    import csv
    import statistics
    runs = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if "k_" in r["Kernel_Name"]:
                runs.setdefault(short_kernel(r["Kernel_Name"]), []).append(
                    int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    out = {}
    for k, v in runs.items():
        v.sort()
        cut = len(v) // 10
        trim = v[cut:len(v) - cut]
        out[k] = {"launches": len(v), "median_us": round(statistics.median(v) / 1e3, 2),
                  "trimmed_mean_us": round(sum(trim) / len(trim) / 1e3, 2), "min_us": round(v[0] / 1e3, 2)}
        log("%-40s %5d launches  median %8.2f us  trimmed mean %8.2f us  min %8.2f us"
            % (k, len(v), out[k]["median_us"], out[k]["trimmed_mean_us"], out[k]["min_us"]))
    return out


def _counter_means(path):
    """{kernel: {counter: mean per dispatch}, kernel: dispatches} from a rocprofv3
    counter collection (CSV). A decompose kernel launched over one operand and
    over both (a and b) is keyed apart: `name@1` and `name@2`."""
    import csv
    vals, disp = {}, {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if "k_" not in r["Kernel_Name"]:
                continue
            k = short_kernel(r["Kernel_Name"])
            if "batch" in k and not k.startswith(("k_ntt_rows", "k_crt")):
                # operands per launch: the grid over the one-operand grid at the bench point
                one = {"k_decompose_cols": N, "k_decompose_mfma": 4 * N, "k_ntt_col": 4096 * 134}
                unit = next((v for pre, v in one.items() if k.startswith(pre)), None)
                if unit:
                    k += "@%d" % max(1, int(r["Grid_Size"]) // unit)
            vals.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
            disp.setdefault(k, set()).add(r["Dispatch_Id"])
    return ({k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in vals.items()},
            {k: len(d) for k, d in disp.items()})


# Algorithmic bytes per launch at the bench point (K = 134, L = 64): what each
# kernel must read and write, for the traffic table's last column.
def _alg_bytes(kernel):
    kn, nl = 4 * 134 * N, 4 * N * 64
    if kernel.startswith(("k_decompose", "k_crt")):
        return (kn + nl) * (int(kernel.split("@")[1]) if "@" in kernel else 1)
    if kernel.startswith("k_ntt_rows_fused"):
        return 3 * kn
    if kernel.startswith("k_ntt_col"):
        return 2 * kn * (int(kernel.split("@")[1]) if "@" in kernel else 1)
    if kernel == "k_copy16":
        return 2 * (256 << 20)
    if kernel == "k_sum16":
        return 256 << 20
    return None


def traffic_table(fetch_csv, write_csv, log):
    """HBM-side bytes per dispatch from a FETCH_SIZE pass and a WRITE_SIZE pass of
    `--phases-only --copy-yardstick` (rocprofv3 counter CSVs). The factors come
    from the yardsticks in the same runs: the copy reads and writes 256 MiB, the
    read-only pass reads 256 MiB."""
    fetch, disp = _counter_means(fetch_csv)
    write, _ = _counter_means(write_csv)
    f_read = (256 << 10) / fetch["k_sum16"]["FETCH_SIZE"]        # KiB read / counter
    f_write = (256 << 10) / write["k_copy16"]["WRITE_SIZE"]
    log("# bytes = (FETCH_SIZE * %.4f + WRITE_SIZE * %.4f) * 1024, mean over the dispatches of each kernel"
        % (f_read, f_write))
    log("# %-40s %12s %11s %11s %11s %12s %9s" % ("kernel", "bytes", "fetch_KiB", "write_KiB", "dispatches",
                                                   "algorithmic", "bytes/alg"))
    out = {}
    for k in fetch:
        fk, wk = fetch[k].get("FETCH_SIZE", 0.0), write.get(k, {}).get("WRITE_SIZE", 0.0)
        b = (fk * f_read + wk * f_write) * 1024
        alg = _alg_bytes(k)
        out[k] = {"bytes": round(b), "fetch_KiB": round(fk, 1), "write_KiB": round(wk, 1), "dispatches": disp[k],
                  "algorithmic": alg}
        log("%-42s %12d %11.1f %11.1f %11d %12s %9s" % (k, b, fk, wk, disp[k], alg if alg else "-",
                                                       "%.3f" % (b / alg) if alg else "-"))
    return out


SQ_COLUMNS = ["SQ_INSTS_VALU", "SQ_INSTS_MFMA", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_SMEM",
              "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR", "SQ_LDS_BANK_CONFLICT", "SQ_WAVE_CYCLES", "SQ_BUSY_CYCLES",
              "SQ_WAIT_INST_ANY", "SQ_WAIT_ANY", "SQ_ACTIVE_INST_LDS", "SQ_VALU_MFMA_BUSY_CYCLES",
              "SQ_ACTIVE_INST_VALU"]


def sq_table(paths, log):
    """SQ counters per wave (counter mean per dispatch / SQ_WAVES) from one or
    more rocprofv3 counter passes over the same run (CSV)."""
    merged = {}
    for path in paths:
        for k, cs in _counter_means(path)[0].items():
            merged.setdefault(k, {}).update(cs)
    log("# %-38s %7s " % ("kernel", "waves") + " ".join("%12s" % c[3:15].lower() for c in SQ_COLUMNS))
    out = {}
    for k, cs in merged.items():
        w = cs.get("SQ_WAVES")
        if not w:
            continue
        out[k] = {"waves": round(w), **{c: round(cs[c] / w, 1) for c in SQ_COLUMNS if c in cs}}
        log("%-40s %7d " % (k, round(w)) + " ".join("%12s" % ("%.1f" % (cs[c] / w) if c in cs else "-")
                                                     for c in SQ_COLUMNS))
    return out


def clock_table(path, log, xcds=8, min_us=200.0):
    """The engine clock each kernel ran at, from a rocprofv3 GRBM_COUNT pass
    (CSV with the dispatches' timestamps): GRBM_COUNT counts engine clocks
    summed over the device's XCDs, so count / XCDs / duration is the clock.
    Only dispatches of at least `min_us` count: the counter's window is a few
    microseconds wider than the dispatch, which skews short ones."""
    import csv
    per = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != "GRBM_COUNT" or "k_" not in r["Kernel_Name"]:
                continue
            ns = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            if ns < min_us * 1e3:
                continue
            per.setdefault(short_kernel(r["Kernel_Name"]), []).append((ns, float(r["Counter_Value"])))
    log("# %-48s %10s %12s %9s" % ("kernel", "dispatches", "mean us", "GHz"))
    out = {}
    for k, v in per.items():
        ghz = sum(c / xcds / ns for ns, c in v) / len(v)
        out[k] = {"dispatches": len(v), "mean_us": round(sum(ns for ns, _ in v) / len(v) / 1e3, 1),
                  "ghz": round(ghz, 3)}
        log("%-50s %10d %12.1f %9.3f" % (k, len(v), out[k]["mean_us"], ghz))
    return out


def trace_shares(path, replays, log):
    """The last of `replays` graph replays in a rocprofv3 kernel trace (CSV,
    columns Kernel_Name, Start_Timestamp, End_Timestamp in ns): every instant
    of its span is split evenly among the kernels running then. Returns per
    kernel (template arguments dropped) its microseconds of the replay, its
    share, and per product (16 per replay)."""
    import csv
    with open(path) as f:   # the library's kernels are named k_*; template arguments dropped
        rows = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short_kernel(r["Kernel_Name"]).split("<")[0])
                for r in csv.DictReader(f) if "k_" in r["Kernel_Name"]]   # runtime copy kernels left out
    rows.sort()
    per = len(rows) // replays
    if per == 0 or len(rows) % replays:
        raise SystemExit("%d dispatches do not split into %d equal replays" % (len(rows), replays))
    last = rows[-per:]
    edges = sorted({t for s, e, _ in last for t in (s, e)})
    share = {}
    idle = 0
    for t0, t1 in zip(edges, edges[1:]):
        run = [k for s, e, k in last if s <= t0 and e >= t1]
        if not run:
            idle += t1 - t0
        for k in run:
            share[k] = share.get(k, 0.0) + (t1 - t0) / len(run)
    span = edges[-1] - edges[0]
    out = {"dispatches_per_replay": per, "replay_us": round(span / 1e3, 2), "idle_us": round(idle / 1e3, 2),
           "kernels": {k: {"us": round(v / 1e3, 2), "share": round(v / span, 4), "per_product_us": round(v / 1e3 / BATCH, 2)}
                       for k, v in sorted(share.items(), key=lambda kv: -kv[1])}}
    for k, v in out["kernels"].items():
        log("%-34s %9.1f us  %5.1f %%  %6.2f us per product" % (k, v["us"], 100 * v["share"], v["per_product_us"]))
    return out


def uneven_rate(args, log):
    """The 64-product batch of uneven shapes (N in 4096 .. 65536, q of 128 .. 2019
    bits, seeds 4000 .. 4063 as in the golden fixtures), device-resident, replayed
    as one hipGraph: products/s and mismatches against the GMP digests. A second
    workload beside the headline, never the headline."""
    import polymult  # noqa: E402
    path = os.path.join(ROOT, "tests", "golden", "golden.json")
    with open(path) as f:
        cases = sorted((c for c in json.load(f)["cases"] if c["name"].startswith("config-uneven-")),
                       key=lambda c: c["name"])
    points = [make_point(c["seed"], c["N"], c["W"]) for c in cases]
    plans = [polymult.Plan(c["N"], c["W"], q, form=polymult.FORM_MFMA) for c, (q, _, _) in zip(cases, points)]
    da = [polymult.DeviceBuffer.from_array(a) for _, a, _ in points]
    db = [polymult.DeviceBuffer.from_array(b) for _, _, b in points]
    dc = [polymult.DeviceBuffer(a.nbytes) for _, a, _ in points]
    t0 = time.perf_counter()
    graph = polymult.Batch(plans, da, db, dc)
    shape = shape_info(graph, time.perf_counter() - t0)
    log("uneven batch shape (timed choice): %s" % shape)
    graph.launch()
    polymult.synchronize()
    mismatches = sum(digest(x.download((c["N"], c["L"]))) != c["c_sha"] for c, x in zip(cases, dc))
    for _ in range(args.warmup):
        graph.launch()
    polymult.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch()
    polymult.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    out = {"value": round(len(cases) / sec, 2), "unit": "polymuls/s", "products": len(cases),
           "ms_per_batch": round(sec * 1e3, 4), "mismatches": int(mismatches), "shape": shape,
           "note": "64 products of uneven N (4096-65536) and q width (128-2019 bits), one hipGraph, device-resident"}
    log("uneven batch: %.3f ms per batch of %d, %.1f products/s, %d mismatches"
        % (sec * 1e3, len(cases), len(cases) / sec, mismatches))
    graph.close()
    for x in da + db + dc:
        x.close()
    for p in plans:
        p.close()
    return out


# The reference's two measured points (its working notes, on an NVIDIA L40S):
# N=65536 with the spec's 1747-bit prime q, and N=32768 with its 868-bit prime q;
# end-to-end from host vectors, one product per call.
REFERENCE_POINTS = [("point1", "matrix-n65536-w1747-prime", 2.30), ("point2", "matrix-n32768-w868-prime", 0.605)]


def driver_times(n, w, q, a, b, c_sha, reps, log):
    """fherma_init / fherma_run / fherma_free through the harness driver
    (fherma-polymult_amd/solution) on a point directory of `reps` cases, as the
    reference's main.cpp times them (main.cpp:180-186): (init seconds, the
    per-case seconds of fherma_run alone, cases whose answer differs from GMP)."""
    import tempfile
    solution = os.path.join(ROOT, "fherma-polymult_amd", "solution")
    with tempfile.TemporaryDirectory(prefix="pm_point_") as root:
        os.makedirs(os.path.join(root, "point"))
        q.astype("<u4").tofile(os.path.join(root, "point", "q.bin"))
        with open(os.path.join(root, "manifest.json"), "w") as f:
            json.dump({"N": n, "W": w, "L": q.shape[0], "cases": reps}, f)
        for i in range(reps):
            d = os.path.join(root, "cases", "%06d" % i)
            os.makedirs(d)
            a.astype("<u4").tofile(os.path.join(d, "a.bin"))
            b.astype("<u4").tofile(os.path.join(d, "b.bin"))
        r = subprocess.run([solution, root], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError("driver failed: " + r.stderr[-2000:])
        with open(os.path.join(root, "out", "results.json")) as f:
            res = json.load(f)
        bad = sum(digest(np.fromfile(os.path.join(root, "out", "%06d" % i, "c.bin"), dtype="<u4")) != c_sha
                  for i in range(reps))
        if any(c["status"] != "ok" for c in res["cases"]):
            raise RuntimeError("driver: a case crashed")
        return res["init_s"], [c["seconds"] for c in res["cases"]], bad


def reference_points(log, reps=11):
    """The reference's measured points with their golden q: the median time of
    fherma_run per product through the harness driver (fherma_run's own path:
    host vectors in, the answer's vector sized and filled, PCIe included; the
    host modes from fherma_init's timed choice), the same through pm_multiply on
    numpy arrays, and the mean device-resident time of one product (no
    batching); every answer checked against its GMP digest."""
    import polymult  # noqa: E402
    path = os.path.join(ROOT, "tests", "golden", "golden.json")
    with open(path) as f:
        cases = {c["name"]: c for c in json.load(f)["cases"]}
    out = {}
    for key, name, ref_ms in REFERENCE_POINTS:
        c = cases[name]
        n, w, L = c["N"], c["W"], c["L"]
        qi = int(c["q_hex"], 16)
        q = np.array([(qi >> (32 * l)) & 0xFFFFFFFF for l in range(L)], dtype=np.uint32)
        _, a, b = make_operands(c["seed"], n, w, q)
        init_s, secs, mism = driver_times(n, w, q, a, b, c["c_sha"], reps, log)
        plan = polymult.Plan(n, w, q, form=polymult.FORM_MFMA)
        mode, _ = plan.tune_host()
        mism += int(digest(plan.multiply(a, b)) != c["c_sha"])
        host = []
        for _ in range(reps):
            t0 = time.perf_counter()
            plan.multiply(a, b)
            host.append(time.perf_counter() - t0)
        da, db = polymult.DeviceBuffer.from_array(a), polymult.DeviceBuffer.from_array(b)
        dc = polymult.DeviceBuffer(a.nbytes)
        plan.multiply_device(da, db, dc)
        polymult.synchronize()
        mism += int(digest(dc.download((n, L))) != c["c_sha"])
        iters = 20
        t0 = time.perf_counter()
        for _ in range(iters):
            plan.multiply_device(da, db, dc)
        polymult.synchronize()
        dev = (time.perf_counter() - t0) / iters
        names = {polymult.HOST_STAGED: "staged", polymult.HOST_MAPPED: "mapped", polymult.HOST_DIRECT: "direct"}
        out[key] = {"N": n, "W": w, "K": plan.K, "q": "the spec's prime", "mismatches": mism,
                    "fherma_run_ms": round(float(np.median(secs)) * 1e3, 4),
                    "fherma_init_s": round(init_s, 3),
                    "pm_multiply_ms": round(float(np.median(host)) * 1e3, 4),
                    "host_modes": "%s in, %s out" % tuple(names.get(m, str(m)) for m in plan.host_modes),
                    "device_resident_ms": round(dev * 1e3, 4),
                    "reference_end_to_end_ms_l40s": ref_ms,
                    "note": "fherma_run_ms: the driver's per-case seconds around fherma_run (median of %d), "
                            "with the answer's vector sized inside the call" % reps}
        log("%s (N=%d, W=%d): fherma_run %.3f ms (driver, median of %d), pm_multiply %.3f ms (%s), "
            "device-resident %.3f ms, %d mismatches"
            % (key, n, w, out[key]["fherma_run_ms"], reps, out[key]["pm_multiply_ms"], out[key]["host_modes"],
               dev * 1e3, mism))
        for x in (da, db, dc):
            x.close()
        plan.close()
    return out


def issue_probe(log):
    """Cycles per wave64 instruction per SIMD for each probe kind of
    pm_probe_issue (fherma-polymult_amd/csrc/pm_probe.hip), at the device's
    reported peak clock: one table line per kind."""
    import polymult  # noqa: E402
    name, cus = polymult.device_info()
    simds = 4 * cus
    rows = []
    for kind, iters in ((0, 4096), (1, 4096), (2, 4096), (3, 2048), (4, 2048), (5, 4096), (6, 4096),
                        (7, 2048), (8, 2048), (9, 1024), (10, 2048), (11, 1024)):
        r = polymult.probe_issue(kind, iters)
        cycles = r["ms"] * 1e-3 * r["clock_mhz"] * 1e6
        n = r["valu_per_simd"] + r["mfma_per_simd"]
        r.update({"cycles_per_simd": round(cycles), "simds": simds,
                  "cycles_per_valu": round(cycles / r["valu_per_simd"], 3) if r["valu_per_simd"] and
                  not r["mfma_per_simd"] else None,
                  "cycles_per_mfma": round(cycles / r["mfma_per_simd"], 3) if r["mfma_per_simd"] and
                  not r["valu_per_simd"] else None})
        log("%-32s %8.3f ms  %12.0f VALU  %10.0f MFMA per SIMD  %10d cycles at %.0f MHz" %
            (r["kind"], r["ms"], r["valu_per_simd"], r["mfma_per_simd"], r["cycles_per_simd"], r["clock_mhz"]))
        rows.append(r)
    return {"device": name, "cus": cus, "probes": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (PCIe-inclusive) rate")
    ap.add_argument("--no-uneven", action="store_true", help="skip the 64-product uneven batch")
    ap.add_argument("--no-points", action="store_true", help="skip the reference's two measured points")
    ap.add_argument("--phase-iters", type=int, default=20,
                    help="back-to-back launches of each phase's kernel in the per-phase timing")
    ap.add_argument("--phases-only", action="store_true",
                    help="only the per-kernel HIP-event profile (one product at a time on one stream), so that "
                         "a kernel trace of this run holds no concurrently running kernels")
    ap.add_argument("--copy-yardstick", action="store_true",
                    help="first run three 256 MiB device copies and three 256 MiB read-only passes (the "
                         "known-size kernels that scale the HBM traffic counters)")
    ap.add_argument("--graph-only", action="store_true",
                    help="only the 16-product batch graph: create it, replay it warmup + steps times, print "
                         "its rate (for profiling the graph itself)")
    ap.add_argument("--trace-stats", metavar="CSV",
                    help="per-kernel duration statistics (count, median, 10 %%-trimmed mean, min) of a "
                         "rocprofv3 kernel trace (CSV)")
    ap.add_argument("--trace-shares", metavar="CSV",
                    help="read a rocprofv3 kernel trace (CSV) of a --graph-only run and print how the wall time "
                         "of its last replay divides among the kernels running at each instant")
    ap.add_argument("--traffic-table", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="HBM bytes per dispatch from a FETCH_SIZE and a WRITE_SIZE rocprofv3 pass of "
                         "--phases-only --copy-yardstick, scaled by the yardsticks")
    ap.add_argument("--sq-table", nargs="+", metavar="CSV",
                    help="SQ counters per wave from rocprofv3 counter passes of one workload")
    ap.add_argument("--clock-table", metavar="CSV",
                    help="the engine clock per kernel from a GRBM_COUNT rocprofv3 pass (dispatches of 200 us "
                         "and more)")
    ap.add_argument("--issue-probe", action="store_true",
                    help="only the issue-rate probes (cycles per instruction kind on each SIMD), as JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the products of a batch are "
              "independent and would be split whole across GPUs)", file=sys.stderr)
        sys.exit(2)

    import polymult  # noqa: E402

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    if args.issue_probe:
        print(json.dumps(issue_probe(log)), flush=True)
        return
    if args.trace_stats:
        print(json.dumps(trace_stats(args.trace_stats, log)), flush=True)
        return
    if args.trace_shares:
        print(json.dumps(trace_shares(args.trace_shares, args.steps + args.warmup + 1, log)), flush=True)
        return
    if args.traffic_table:
        print(json.dumps(traffic_table(*args.traffic_table, log)), flush=True)
        return
    if args.sq_table:
        print(json.dumps(sq_table(args.sq_table, log)), flush=True)
        return
    if args.clock_table:
        print(json.dumps(clock_table(args.clock_table, log)), flush=True)
        return

    if args.copy_yardstick:
        # known-size yardsticks for the traffic counters: a 256 MiB copy (reads and
        # writes 256 MiB) three times, then a read-only pass over 256 MiB three times
        src, dst = polymult.DeviceBuffer(256 << 20), polymult.DeviceBuffer(256 << 20)
        for _ in range(3):
            polymult.device_copy(dst, src)
        polymult.synchronize()
        for _ in range(3):
            polymult.device_sum(src)
        src.close()
        dst.close()

    log("generating %d points" % BATCH)
    points = [make_point(SEED0 + i, N, W) for i in range(BATCH)]
    plans = [polymult.Plan(N, W, q, form=polymult.FORM_MFMA) for q, _, _ in points]
    K, L = plans[0].K, plans[0].L
    da = [polymult.DeviceBuffer.from_array(a) for _, a, _ in points]
    db = [polymult.DeviceBuffer.from_array(b) for _, _, b in points]
    dc = [polymult.DeviceBuffer(4 * N * L) for _ in points]

    gold = golden_digests()
    value = sec = plain_value = shape = None
    if args.graph_only:
        args.no_plain = True
        value, sec, _, mismatches, shape = batch_rates(args, plans, da, db, dc, gold, log)
        print(json.dumps({"graph_only": True, "value": round(value, 2), "ms_per_step": round(sec * 1e3, 4),
                          "replays": args.warmup + args.steps + 1, "mismatches": int(mismatches),
                          "shape": shape}), flush=True)
        return
    if args.phases_only:
        for p, a, b, c in zip(plans, da, db, dc):
            p.multiply_device(a, b, c)
        polymult.synchronize()
        mismatches = sum(digest(dc[i].download((N, L))) != gold.get(SEED0 + i) for i in range(BATCH))
        log("one product at a time: %d mismatches vs GMP" % mismatches)
    else:
        value, sec, plain_value, mismatches, shape = batch_rates(args, plans, da, db, dc, gold, log)
    uneven = None
    if not (args.phases_only or args.no_uneven):
        uneven = uneven_rate(args, log)
    ref_points = None
    if not (args.phases_only or args.no_points):
        ref_points = reference_points(log)

    # Host buffers in and out (fherma_run's path: PCIe transfers included), after
    # the timed choice of the host mode on the first plan: the 16 products one
    # after another, and as a host batch with 2-4 products in flight at once
    # (pm_batch_multiply_host). Never the headline value.
    host = None
    if not (args.no_host or args.phases_only):
        plans[0].tune_host()
        modes = plans[0].host_modes
        for p in plans[1:]:
            p.set_host_modes(*modes)
        names = {polymult.HOST_STAGED: "staged", polymult.HOST_MAPPED: "mapped", polymult.HOST_DIRECT: "direct"}
        mode = "%s in, %s out" % tuple(names.get(m, str(m)) for m in modes)
        A = [a for _, a, _ in points]
        B = [b for _, _, b in points]
        outs = [np.empty((N, L), dtype=np.uint32) for _ in points]
        per_depth = {}
        host_bad = 0
        for depth in (1, 2, 3, 4):
            polymult.batch_multiply_host(plans, A, B, depth=depth, out=outs)
            host_bad += sum(digest(outs[i]) != gold.get(SEED0 + i) for i in range(BATCH))
            t0 = time.perf_counter()
            for _ in range(2):
                polymult.batch_multiply_host(plans, A, B, depth=depth, out=outs)
            per_depth[depth] = 2 * BATCH / (time.perf_counter() - t0)
            log("host buffers, %d in flight: %.3f ms per product (%s)" % (depth, 1e3 / per_depth[depth], mode))
        best = max(per_depth, key=per_depth.get)
        host = {"value": round(per_depth[best], 2), "unit": "polymuls/s", "mode": mode, "in_flight": best,
                "one_at_a_time": round(per_depth[1], 2),
                "per_in_flight": {str(d): round(v, 2) for d, v in per_depth.items()},
                "mismatches": int(host_bad),
                "note": "numpy arrays in and out (PCIe included): the 16 products through the plans' host "
                        "pipelines, `in_flight` at once (pm_batch_multiply_host)"}
        mismatches += host_bad

    # per-phase HIP-event timings on the plan's stream (mean over the batch)
    phases = {}
    for p, a, b, c in zip(plans, da, db, dc):
        for k, v in p.profile_phases(a, b, c, args.phase_iters).items():
            phases[k] = phases.get(k, 0.0) + v / BATCH
    # phases a plan's schedule leaves empty (reported as 0: the forward column
    # passes, when their stages run inside decompose and the row pass)
    phases = {k: v for k, v in phases.items() if v > 0}
    log("phases (ms): " + ", ".join("%s %.4f" % kv for kv in phases.items()))
    nbytes = phase_bytes(K, L)
    dom = max(phases, key=phases.get)
    achieved = nbytes[dom] / (phases[dom] * 1e-3) / 1e9
    kname = phase_kernels(plans[0])[dom]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kname), "kernel": dom,
                "kernel_name": kname, "kernel_ms": round(phases[dom], 5),
                "algorithmic_bytes": nbytes[dom]}
    prod_bytes = 12 * N * L + 56 * K * N
    # the bytes a product of the measured schedule must move: the two decompositions
    # 8 N L + 8 K N and the CRT 4 N L + 4 K N, the fused row pass 12 K N and the
    # inverse column pass 8 K N; the reference's schedule adds two forward column
    # passes, 16 K N
    sched_bytes = 12 * N * L + (32 if plans[0].fused_cols == 1 else 48) * K * N
    cpu = None if args.no_cpu else cpu_baseline(points[:1])
    name, cus = polymult.device_info()
    roofline["valu_issue"] = valu_issue(kname, phases[dom], cus)
    out = {
        "metric": "polymuls/s at N=65536, 2048-bit q: batched+MFMA vs plain HIP; % HBM roofline",
        "value": None if value is None else round(value, 2), "unit": "polymuls/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": None if sec is None else round(sec * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": "batch of 16 products, N=65536, random odd 2048-bit q (L=64), device-resident", "form": "int8-MFMA decompose/CRT, batch replayed as one hipGraph", "batch": BATCH, "batch_shape": shape, "N": N, "W": W, "L": L, "K": K,
                   "parallelism": "single GPU", "device": name, "cus": cus,
                   "row_twiddles": "generated" if plans[0].twiddles == polymult.TWIDDLES_GEN else "table",
                   "prime_basis": "p < 2^31 (the reference's rule)"},
        "plain_value": None if plain_value is None else round(plain_value, 2),
        "mfma_over_plain": None if plain_value is None else round(value / plain_value, 3),
        "mismatches": int(mismatches),
        "effective_GBps_per_product_yardstick": None if value is None else round(prod_bytes * value / 1e9, 1),
        "phases_ms": {k: round(v, 5) for k, v in phases.items()},
        "host_inclusive": host,
        "uneven_batch": uneven,
        "reference_points": ref_points,
        "roofline": roofline,
        # the whole product of the measured batch against the same roofline: the
        # schedule's algorithmic bytes per product over the batch's time per product
        "roofline_product": None if value is None else {
            "algorithmic_bytes": sched_bytes, "us_per_product": round(1e6 / value, 2),
            "achieved_GBps": round(sched_bytes * value / 1e9, 1),
            "frac": round(sched_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
            "schedule": "fused column stages" if plans[0].fused_cols == 1 else "the reference's passes"},
        # every phase against the same roofline (the CRT and the decompositions take a larger
        # share of the batch graph's wall time than the fused row pass, which is slowest alone)
        "roofline_by_phase": {k: {"algorithmic_bytes": nbytes[k], "ms": round(v, 5),
                                  "achieved_GBps": round(nbytes[k] / (v * 1e-3) / 1e9, 1),
                                  "frac": round(nbytes[k] / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                              for k, v in phases.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
