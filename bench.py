"""Benchmark of the streaming hot path on one MI355X.

Default workload (one GPU): the full FP64 streaming suite --
BS1-BS5 on N = 2^27 doubles (U[-1,1), LCG seed 1, alpha = 0.37, beta = -1.21),
BS6-BS8 on a 64^3-element hex box at degree 7 (134,217,728 local nodes,
90,518,849 global nodes, same seed), BS0 as 10^4 empty launches, plain and
replayed from a HIP graph.

One timed "step" is one BS8 gather+scatter (ogs_t::GatherScatter(q, 1, Add,
Sym)) over the whole mesh; the headline value is its algorithmic HBM
bandwidth, 4(Ng+1) + 4Nl + 16Nl bytes per call (bs8Run.cpp:56-63), over the
host-timed K steps.  Every other kernel is timed the same way and reported
under "suite".  The roofline object uses the BS8 kernel's average duration
from HIP events recorded on the library's stream.  The CPU baseline times
the oracle's restatement of the reference's Serial mode on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import glob
import hashlib
import json
import math
import os
import platform as pyplatform
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SPEC_HBM_GBS = 8000.0
METRIC = "HBM GB/s (% peak) + GDOF/s per BS kernel; mixed/batched vs plain FP64"
PROFILE_COUNTERS = os.path.join(ROOT, "profiles", "bs8_gather_scatter_64cube_p7_counters.md")
BIN = os.path.join(ROOT, "streamparanumal_amd", "bin")

# kernel-name patterns of the suite's dominant kernels, for the counter passes
KERNEL_PATTERNS = {
    "BS1_copy": r"k_copy2<4, true>",
    "BS2_axpy": r"k_axpy2<",
    "BS3_norm_reference": r"k_red_ref1<0, \d+, \d+, \w+, false>",
    "BS4_dot_reference": r"k_red_ref1<1, \d+, \d+, \w+, false>",
    "BS5_cg_update_reference": r"k_red_ref1<2, \d+, \d+, \w+, false>",
    "BS5_cg_full_reference": r"k_red_ref1<3, \d+, \d+, \w+, false>",
    "BS3_norm_fast": r"k_red_ref1<0, \d+, \d+, \w+, true>",
    "BS4_dot_fast": r"k_red_ref1<1, \d+, \d+, \w+, true>",
    "BS5_cg_update_fast": r"k_red_ref1<2, \d+, \d+, \w+, true>",
    "BS5_cg_full_fast": r"k_red_ref1<3, \d+, \d+, \w+, true>",
    "BS5_cg_full_f32": r"k_cg_mixed1<float",
    "BS5_cg_full_bf16": r"k_cg_mixed1<[^>]*bf16_t",
    "BS6_gather": r"k_gather_srb<double",
    "BS7_scatter": r"k_scatter_srb<double",
    "BS8_gather_scatter": r"k_gs_srb<double",
    "BS8_gather_scatter_f32": r"k_gs_srb<float",
    "BS8_gather_scatter_bf16": r"k_gs_srb<[^,]*bf16_t",
}


def source_hash():
    """16 hex digits of sha256 over the sources that decide what the kernels do and move: the HIP
    kernels, the host layer that builds their layouts, and the C headers."""
    pkg = os.path.join(ROOT, "streamparanumal_amd")
    files = sorted(glob.glob(os.path.join(pkg, "csrc", "*")) + glob.glob(os.path.join(pkg, "libp", "src", "*.cpp")) +
                   glob.glob(os.path.join(pkg, "libp", "include", "**", "*.hpp"), recursive=True) +
                   glob.glob(os.path.join(ROOT, "include", "*.h")))
    h = hashlib.sha256()
    for f in files:
        h.update(os.path.relpath(f, ROOT).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


SETTLE_SECONDS = 0.25


def timed(fn, steps, warmup, settle=SETTLE_SECONDS):
    """(host seconds per call, event seconds per call) over `steps` calls after `warmup` calls.

    Before the warm-up calls, `fn` also runs untimed for `settle` seconds of wall time, so the
    timed calls see the GPU at its sustained clocks rather than ramping up from idle after the
    host-side setup (a few warm-up calls of a 0.5 ms kernel are over before the clocks settle)."""
    from streamparanumal_amd import device
    t_end = time.perf_counter() + settle
    while time.perf_counter() < t_end:
        for _ in range(8):
            fn()
        device.finish()
    for _ in range(warmup):
        fn()
    device.finish()
    e0, e1 = device.Event(), device.Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    device.finish()
    host = (time.perf_counter() - t0) / steps
    return host, e0.elapsed_ms(e1) * 1e-3 / steps


def read_traffic():
    """(HBM bytes per BS8 launch, provenance) from the committed rocprofv3 counter summary, or
    (None, reason) when it is missing or was made from other kernel sources than these."""
    if not os.path.exists(PROFILE_COUNTERS):
        return None, "no committed counter profile"
    text = open(PROFILE_COUNTERS).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    h = re.search(r"source_sha16:\s*([0-9a-f]{16})", text)
    if not m:
        return None, "no traffic figure in the committed profile"
    if not h or h.group(1) != source_hash():
        return None, "committed profile is from other kernel sources (stale): not used"
    return float(m.group(1)), f"profiles/{os.path.basename(PROFILE_COUNTERS)} (same kernel sources)"


def traffic_probe(nx, p, log2n, reps=3, batched=False):
    """The launches the counter passes measure (run as the profiled child process): the plain
    calibration copy of 2^27 doubles, then `reps` calls of every suite kernel on the bench's inputs
    (batched: instead, `reps` steps of the 64-mesh batch in FP32 and in FP64 with the default
    launch shapes)."""
    import numpy as np

    import streamparanumal_amd as spx
    from streamparanumal_amd import device, ogs, streaming
    from streamparanumal_amd.device import DeviceArray

    ogs.platform_open(0)
    if batched:
        from streamparanumal_amd.batch import Batch, instances
        c0, c1 = DeviceArray(1 << 27), DeviceArray(1 << 27)
        device.fill_uniform(c0, 1, 0)
        device.set_tuning("bs1", unroll=4, nontemporal=0)
        for _ in range(reps):
            streaming.copy(c0, c1)
        device.set_tuning("bs1")
        del c0, c1
        b = Batch(instances(64, 4), 3, np.float32)
        for bb in (b, b.sibling(np.float64)):
            bb.allocate(seed=5)
            bb.make_plan()
            for _ in range(reps):
                bb.step(0.37)
            device.finish()
            bb.vectors = None
        device.check_kernel_errors()
        return
    n = 1 << log2n
    a, b, pv, xv = (DeviceArray(n) for _ in range(4))
    for i, v in enumerate((a, b, pv, xv)):
        device.fill_uniform(v, 1, i * n)
    c0, c1 = DeviceArray(1 << 27), DeviceArray(1 << 27)
    device.fill_uniform(c0, 1, 0)
    device.set_tuning("bs1", unroll=4, nontemporal=0)
    for _ in range(reps):
        streaming.copy(c0, c1)                       # k_copy2<4, false>: 1 GiB each way
    device.set_tuning("bs1")
    del c0, c1
    for _ in range(reps):
        streaming.copy(a, b)
    for _ in range(reps):
        streaming.axpy(0.37, a, -1.21, b)
    device.fill_uniform(b, 1, n)
    for mode in (spx.REDUCE_REFERENCE, spx.REDUCE_FAST):
        red = streaming.Reducer(n, mode, 256)
        for _ in range(reps):
            red.norm(a)
        for _ in range(reps):
            red.dot(a, b)
        for _ in range(reps):
            red.cg_update(a, 1e-9, b)
        for _ in range(reps):
            red.cg_full(1e-9, pv, a, xv, b)
    mix = streaming.MixedCG()
    for dt in (np.float32, np.uint16):
        vs = [DeviceArray(n, dt) for _ in range(4)]
        for v, src in zip(vs, (pv, a, xv, b)):
            device.convert(src, v)
        for _ in range(reps):
            mix.cg_full(1e-9, *vs)
        del vs
    del a, b, pv, xv
    g = ogs.Ogs.hex(nx, nx, nx, p, keep_ids=False)
    q, gq = DeviceArray(g.N), DeviceArray(g.NgatherGlobal)
    device.fill_uniform(q, 1, 0)
    for _ in range(reps):
        g.gather_scatter(q)
    for _ in range(reps):
        g.gather(gq, q, 1, spx.ADD, spx.TRANS)
    for _ in range(reps):
        g.scatter(q, gq, 1, spx.NOTRANS)
    for dt in (np.float32, np.uint16):
        qs = DeviceArray(g.N, dt)
        device.convert(q, qs)
        for _ in range(reps):
            g.gather_scatter_mixed(qs)
        del qs
    device.finish()
    device.check_kernel_errors()


def counter_suite(args):
    """FETCH_SIZE / WRITE_SIZE of every suite kernel, measured in this run by two rocprofv3 passes
    over `bench.py --traffic-probe`: {entry: {"read", "write", "traffic", "dispatches"}}."""
    from streamparanumal_amd import counters
    if counters.rocprofv3() is None:
        return None, "rocprofv3 not found"
    cmd = [sys.executable, os.path.abspath(__file__), "--traffic-probe", "--nx", str(args.nx), "--degree",
           str(args.degree), "--log2n", str(args.log2n)]
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    try:
        out = counters.measure(cmd, KERNEL_PATTERNS, timeout=240, env=env)
    except Exception as e:  # the roofline then falls back to the committed profile
        return None, f"counter passes failed: {str(e)[-400:]}"
    return out, "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes in this run"


def run_driver(name, *args, timeout=300):
    """Run bin/<name> with `args` and return its printed rows, {row name: [numbers]}."""
    r = subprocess.run([os.path.join(BIN, name), *args], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{name} {' '.join(args)} failed ({r.returncode}): {(r.stderr or r.stdout)[-800:]}")
    rows = {}
    for ln in r.stdout.splitlines():
        m = re.match(r"\s*(\w+) = \[([^\]]*)\]", ln)
        if m:
            rows[m.group(1)] = [float(x) for x in re.split(r"[,;\s]+", m.group(2).strip()) if x]
    return rows


def driver_suite():
    """Rows of the C++ drivers: BS0 as 10^4 empty launches, plain and replayed from a graph, timed on
    the host after a finish as bs0Run.cpp:70-77 does; BS6-BS8 at 8^3 p=7 (seed 0) and 64^3 p=7
    (seed 1) with three timings per call: back-to-back calls (the BSn row), kernel-only events, and
    a host finish after every call, which is what the reference's ogs_t does inside each call; each
    after the same untimed settle as the bench's own timings."""
    out = {}
    plain = run_driver("BS0", "-n", "10000")["BS0"]
    graph = run_driver("BS0", "-n", "10000", "--graph", "TRUE")["BS0"]
    out["BS0_launch"] = {"launches": int(plain[0]), "plain_us_per_launch": round(plain[1] * 1e6, 3),
                         "graph_us_per_launch": round(graph[1] * 1e6, 3), "timer": "C++ driver bin/BS0"}
    for tag, nx, seed in (("8^3 p=7", 8, 0), ("64^3 p=7", 64, 1)):
        for n, what in (("BS6", "gather"), ("BS7", "scatter"), ("BS8", "gather_scatter")):
            rows = run_driver(n, "-nx", str(nx), "-ny", str(nx), "-nz", str(nx), "-p", "7", "--seed", str(seed),
                              "--settle", str(SETTLE_SECONDS))
            b2b, dev, syn = rows[n], rows[n + "device"], rows[n + "sync"]
            dofs = b2b[1]
            out[f"{n}_{what}_{nx}cube_p7"] = {
                "config": tag, "dofs": int(dofs),
                "back_to_back_us": round(b2b[2] * 1e6, 2), "back_to_back_GBs": b2b[4],
                "kernel_us": round(dev[2] * 1e6, 2), "kernel_GBs": dev[4],
                "synced_us": round(syn[2] * 1e6, 2), "synced_GBs": syn[4], "synced_GDOFs": round(syn[3] / 1e9, 3)}
            log("driver", n, tag, json.dumps(out[f"{n}_{what}_{nx}cube_p7"]))
    return out


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return pyplatform.processor() or "unknown"


def cpu_rows(c1_gpu_checksum=None):
    """The reference's Serial and OpenMP modes restated by the oracle, timed on this host:
    BASELINE config 1 (BS8 on 8^3 p=7, seed 0, Serial) with the checksum of its output; BS1-BS5
    Serial on N = 2^24 and OpenMP on N = 2^27; BS6-BS8 Serial and OpenMP on a 32^3 p=7 box."""
    from oracle import pyoracle
    pyoracle.lib()
    T = pyoracle.max_threads()
    rows = {"cpu_model": cpu_model(), "threads_available": T}
    c1 = pyoracle.time_ogs(8, 7, 200, [1])
    nl, ng = c1["nlocal"], c1["nglobal"]
    t = c1["seconds"][1][2]
    rows["C1_BS8_serial"] = {"ms_per_call": round(t * 1e3, 4), "GBs": round((4 * (ng + 1) + 20 * nl) / t / 1e9, 3),
                             "GDOFs": round(ng / t / 1e9, 4), "checksum": c1["checksums"][2],
                             "gpu_checksum": c1_gpu_checksum, "cores": 1,
                             "sample": "8^3 hex p=7, U[-1,1) seed 0, 200 calls; checksum = long-double sum of one "
                                       "gather+scatter's output"}
    bytes_per = {1: 16, 2: 24, 3: 8, 4: 16, 5: 24}
    for threads, log2n, reps in ((1, 24, 3), (T, 27, 5)):
        n = 1 << log2n
        mode = "serial" if threads == 1 else "openmp"
        for w, name in ((1, "BS1_copy"), (2, "BS2_axpy"), (3, "BS3_norm"), (4, "BS4_dot"), (5, "BS5_cg_update")):
            sec = pyoracle.time_stream(w, n, reps, threads)
            rows[f"{name}_{mode}"] = {"N": n, "ms_per_call": round(sec * 1e3, 3),
                                      "GBs": round(bytes_per[w] * n / sec / 1e9, 2), "cores": threads}
    r = pyoracle.time_ogs(32, 7, 3, [1, T])
    nl, ng = r["nlocal"], r["nglobal"]
    b = {"BS6_gather": 4 * (ng + 1) + 12 * nl + 8 * ng, "BS7_scatter": 4 * (ng + 1) + 12 * nl + 8 * ng,
         "BS8_gather_scatter": 4 * (ng + 1) + 20 * nl}
    for threads in (1, T):
        mode = "serial" if threads == 1 else "openmp"
        for k, name in enumerate(("BS6_gather", "BS7_scatter", "BS8_gather_scatter")):
            sec = r["seconds"][threads][k]
            rows[f"{name}_{mode}"] = {"box": "32^3 p=7", "ms_per_call": round(sec * 1e3, 3),
                                      "GBs": round(b[name] / sec / 1e9, 2), "GDOFs": round(ng / sec / 1e9, 4),
                                      "cores": threads}
    for k, v in rows.items():
        log("cpu", k, json.dumps(v))
    return rows


def cpu_baseline():
    """The oracle's Serial restatement of BS8 on a 32^3 p=7 box, 1 host thread (the JSON line's
    cpu_baseline), and the same row loop under OpenMP on every host thread the process may use
    (the reference's OpenMP mode; reported in the suite)."""
    from oracle import pyoracle
    pyoracle.lib()
    nx, p, reps = 32, 7, 30
    out = {}
    for threads in (1, pyoracle.max_threads()):
        r = pyoracle.time_bs8(nx, p, reps, threads)
        nl, ng = r["nlocal"], r["nglobal"]
        bytes_ = 4 * (ng + 1) + 4 * nl + 16 * nl
        mode = "Serial-mode restatement" if threads == 1 else "OpenMP row loop"
        out[threads] = {"value": round(bytes_ / r["seconds"] / 1e9, 3), "unit": "GB/s", "cores": threads,
                        "kind": "port",
                        "sample": f"BS8 gather+scatter FP64, {nx}^3 hex p={p} ({nl} local / {ng} global nodes), "
                                  f"{reps} calls of the oracle's {mode} (gcc -O3), {threads} thread(s); "
                                  f"{r['seconds'] * 1e3:.2f} ms/call, {ng / r['seconds'] / 1e9:.4f} GDOF/s"}
    return out[1], out[max(out)]


def batched_suite(K, W, count=64, seed=4, fields=3):
    """64 independent hex meshes (per-side elements U{4..48}, p in {3,5,7,9}, from the LCG stream of
    `seed`), `fields` fields each: FP32-stored gather+scatter (FP64 row sums) then the CG update
    with an FP64 rdotr per mesh, as one captured graph with autotuned launch shapes; against the
    same work as 64 separate plain FP64 launch sequences, and as the batched graph in FP64."""
    import numpy as np

    import streamparanumal_amd as spx
    from streamparanumal_amd import device, streaming
    from streamparanumal_amd._lib import check, libspx
    from streamparanumal_amd.batch import Batch, SeparateFP64, instances
    from streamparanumal_amd.device import DeviceArray

    alpha = 0.37
    t0 = time.perf_counter()
    meshes = instances(count, seed)
    b = Batch(meshes, fields, np.float32)
    setup_s = time.perf_counter() - t0
    dofs = fields * b.global_dofs
    log(f"batched: {count} meshes, {b.N} local / {b.global_dofs} global nodes, setup {setup_s:.1f}s")
    b.allocate(seed=5)
    t0 = time.perf_counter()
    tune = b.autotune(reps=3)
    log("batched autotune", f"{time.perf_counter() - t0:.1f}s", json.dumps(tune))

    # error of one step against the plain FP64 path (bit-exact with the oracle), from the same inputs
    b.allocate(seed=5)
    sep = SeparateFP64(b, seed=5)
    b.step(alpha)
    sep.step(alpha)
    device.finish()
    nmax = max(fields * g.N for g in b.parts)
    tmp = DeviceArray(nmax)
    red = streaming.Reducer(nmax, spx.REDUCE_FAST, 256)

    def sq(ptr, n):
        check(libspx().spx_bs3_norm(n, ptr, red.tmp.ptr, red.out.ptr, spx.REDUCE_FAST, 256), "spx_bs3_norm")
        return red.value()

    worst = {"q": 0.0, "x": 0.0, "r": 0.0, "rdotr": 0.0}
    rd32 = b.plan.values()
    rd64 = sep.rdotr()
    for i, g in enumerate(b.parts):
        n, off = fields * g.N, int(b.seg_starts[i])
        for key in ("q", "x", "r"):
            ref = sep.vecs[i][key]
            check(libspx().spx_convert_f32_to_f64(n, b.vectors[key].offset(off), tmp.ptr), "convert")
            nref = sq(ref.ptr, n)
            check(libspx().spx_bs2_axpy(n, 1.0, ref.ptr, -1.0, tmp.ptr), "axpy")
            worst[key] = max(worst[key], math.sqrt(sq(tmp.ptr, n) / nref))
        worst["rdotr"] = max(worst["rdotr"], abs(rd32[i] - rd64[i]) / rd64[i])
    del tmp, red
    log("batched error vs FP64 (max over meshes)", json.dumps(worst))

    graph = b.capture(alpha)
    h32, e32 = timed(lambda: graph.launch(), K, W)
    h64s, e64s = timed(lambda: sep.step(alpha), K, W)
    del sep, graph
    b64 = b.sibling(np.float64)
    b64.allocate(seed=5)
    tune64 = b64.autotune(reps=3)
    graph64 = b64.capture(alpha)
    h64b, e64b = timed(lambda: graph64.launch(), K, W)
    del graph64, b64
    out = {"meshes": count, "fields": fields, "local_nodes": b.N, "global_nodes": b.global_dofs,
           "setup_s": round(setup_s, 1),
           "autotuned": tune, "autotuned_f64": tune64,
           "batched_f32_graph": {"ms": round(h32 * 1e3, 3), "kernel_ms": round(e32 * 1e3, 3),
                                 "GDOFs": round(dofs / h32 / 1e9, 2)},
           "batched_f64_graph": {"ms": round(h64b * 1e3, 3), "GDOFs": round(dofs / h64b / 1e9, 2)},
           "separate_f64": {"launches": 3 * count, "ms": round(h64s * 1e3, 3),
                            "GDOFs": round(dofs / h64s / 1e9, 2)},
           "speedup_f32_batched_vs_separate_f64": round(h64s / h32, 3),
           "rel_err_vs_f64_max": {k: float(f"{v:.3e}") for k, v in worst.items()}}
    log("batched", json.dumps(out))
    del b
    return out


def precision_errors(alpha=0.31, seed=3):
    """Errors of the FP32 / BF16 storage forms against this library's FP64 path (bit-exact with the
    oracle) for the U[-1,1) and log-uniform |x| in [1e-6, 1] input sets of `seed`: BS8 on the 8^3
    and 64^3 p=7 boxes, the BS5 CG update (x += alpha p, r -= alpha Ap, rdotr) at N = 262,144 and
    2^27.  Relative errors are ||approx - exact|| / ||exact|| over the whole vector."""
    import numpy as np

    import streamparanumal_amd as spx
    from streamparanumal_amd import device, ogs, streaming
    from streamparanumal_amd.device import DeviceArray

    fills = {"uniform": device.fill_uniform, "loguniform": device.fill_loguniform}
    stores = ((spx.STORE_F32, np.float32, "f32"), (spx.STORE_BF16, np.uint16, "bf16"))
    rows = []

    def rel(red, approx, exact, diff):
        red.norm(exact)
        nref = red.value()
        streaming.copy(approx, diff)
        streaming.axpy(1.0, exact, -1.0, diff)
        red.norm(diff)
        return math.sqrt(red.value() / nref)

    for nx, n5 in ((8, 262144), (64, 1 << 27)):
        g = ogs.Ogs.hex(nx, nx, nx, 7, keep_ids=False)
        Nl = g.N
        red = streaming.Reducer(max(Nl, n5), spx.REDUCE_FAST, 256)
        q, ref, back, diff = (DeviceArray(Nl) for _ in range(4))
        for name, fill in fills.items():
            fill(q, seed, 0)
            streaming.copy(q, ref)
            g.gather_scatter(ref)
            for st, dt, tag in stores:
                qs = DeviceArray(Nl, dt)
                device.convert(q, qs)
                g.gather_scatter_mixed(qs)
                device.convert(qs, back)
                rows.append({"kernel": "BS8", "size": f"{nx}^3 p=7", "inputs": name, "storage": tag,
                             "rel_err": float(f"{rel(red, back, ref, diff):.3e}")})
                del qs
        del q, ref, back, diff, g
        vec = [DeviceArray(n5) for _ in range(4)]          # p, Ap, x, r
        x0, r0, back, diff = (DeviceArray(n5) for _ in range(4))
        red_ref = streaming.Reducer(n5, spx.REDUCE_REFERENCE, 256)
        mix = streaming.MixedCG()
        for name, fill in fills.items():
            for i, v in enumerate(vec):
                fill(v, seed, i * n5)
            streaming.copy(vec[2], x0)
            streaming.copy(vec[3], r0)
            red_ref.cg_full(alpha, vec[0], vec[1], x0, r0)
            rd64 = red_ref.value()
            for st, dt, tag in stores:
                vs = [DeviceArray(n5, dt) for _ in range(4)]
                for v, src in zip(vs, vec):
                    device.convert(src, v)
                mix.cg_full(alpha, *vs)
                rd = mix.value()
                device.convert(vs[2], back)
                ex = rel(red, back, x0, diff)
                device.convert(vs[3], back)
                er = rel(red, back, r0, diff)
                rows.append({"kernel": "BS5 CG update", "size": f"N={n5}", "inputs": name, "storage": tag,
                             "rel_err_x": float(f"{ex:.3e}"), "rel_err_r": float(f"{er:.3e}"),
                             "rel_err_rdotr": float(f"{abs(rd - rd64) / rd64:.3e}")})
                del vs
        del vec, x0, r0, back, diff, red, red_ref, mix
    for r in rows:
        log("precision", json.dumps(r))
    return rows


def element_types(K, W, seed=2):
    """FP64 gather (Trans), scatter (NoTrans) and gather+scatter (Sym) on a box of each element type
    at p=3 and p=7, sized to about 1e8 local nodes, set up as the mesh does (Ogs.box: the general
    ogs setup from the box ids, the hex box in closed form, and the mesh's gather+scatter tile
    order)."""
    import streamparanumal_amd as spx
    from streamparanumal_amd import device, ogs
    from streamparanumal_amd.device import DeviceArray

    rows = []
    for el in ("Tri", "Quad", "Tet", "Hex"):
        for p in (3, 7):
            _, npe = ogs.box_shape(el, 1, 1, 1, p)
            per_cell = {"Tri": 2, "Quad": 1, "Tet": 6, "Hex": 1}[el] * npe
            dim = 2 if el in ("Tri", "Quad") else 3
            n = int(round((1e8 / per_cell) ** (1.0 / dim)))
            nz = n if dim == 3 else 1
            t0 = time.perf_counter()
            g = ogs.Ogs.box(el, n, n, nz, p, keep_ids=False)
            setup = time.perf_counter() - t0
            Nl, Ng = g.N, g.NgatherGlobal
            q = DeviceArray(Nl)
            device.fill_uniform(q, seed, 0)
            gq = DeviceArray(Ng)
            b6 = 4 * (Ng + 1) + 12 * Nl + 8 * Ng
            b7 = 4 * (Ng + 1) + 4 * Nl + 8 * Ng + 8 * Nl
            b8 = 4 * (Ng + 1) + 4 * Nl + 16 * Nl
            _, e6 = timed(lambda: g.gather(gq, q, 1, spx.ADD, spx.TRANS), K, W)
            _, e7 = timed(lambda: g.scatter(q, gq, 1, spx.NOTRANS), K, W)
            _, e8 = timed(lambda: g.gather_scatter(q), K, W)
            device.check_kernel_errors()
            row = {"element": el, "p": p, "n": n, "local_nodes": Nl, "global_nodes": Ng,
                   "Ng_over_Nl": round(Ng / Nl, 3), "setup_s": round(setup, 2),
                   "gather_GBs": round(b6 / e6 / 1e9, 1), "scatter_GBs": round(b7 / e7 / 1e9, 1),
                   "gather_scatter_GBs": round(b8 / e8 / 1e9, 1), "gather_scatter_GDOFs": round(Ng / e8 / 1e9, 2)}
            rows.append(row)
            log("element types", json.dumps(row))
            del g, q, gq
    return rows


def degree_sweep(K, W, seed=2, target=1e8):
    """FP64 gather (Trans), scatter (NoTrans) and gather+scatter (Sym) on hex boxes of
    round((target/(p+1)^3)^(1/3)) elements per side for p = 1..15 (about `target` local nodes
    each), U[-1,1) values from `seed`; library-default launch shapes."""
    import streamparanumal_amd as spx
    from streamparanumal_amd import device, ogs
    from streamparanumal_amd.device import DeviceArray

    rows = []
    for p in range(1, 16):
        n = int(round((target / (p + 1) ** 3) ** (1.0 / 3.0)))
        g = ogs.Ogs.hex(n, n, n, p, keep_ids=False)
        Nl, Ng = g.N, g.NgatherGlobal
        q = DeviceArray(Nl)
        device.fill_uniform(q, seed, 0)
        gq = DeviceArray(Ng)
        b6 = 4 * (Ng + 1) + 12 * Nl + 8 * Ng
        b7 = 4 * (Ng + 1) + 4 * Nl + 8 * Ng + 8 * Nl
        b8 = 4 * (Ng + 1) + 4 * Nl + 16 * Nl
        _, e6 = timed(lambda: g.gather(gq, q, 1, spx.ADD, spx.TRANS), K, W)
        _, e7 = timed(lambda: g.scatter(q, gq, 1, spx.NOTRANS), K, W)
        _, e8 = timed(lambda: g.gather_scatter(q), K, W)
        device.check_kernel_errors()
        row = {"p": p, "n": n, "local_nodes": Nl, "global_nodes": Ng, "Ng_over_Nl": round(Ng / Nl, 3),
               "gather_GBs": round(b6 / e6 / 1e9, 1), "scatter_GBs": round(b7 / e7 / 1e9, 1),
               "gather_scatter_GBs": round(b8 / e8 / 1e9, 1), "gather_GDOFs": round(Ng / e6 / 1e9, 2),
               "scatter_GDOFs": round(Ng / e7 / 1e9, 2), "gather_scatter_GDOFs": round(Ng / e8 / 1e9, 2)}
        rows.append(row)
        log("degree sweep", json.dumps(row))
        del g, q, gq
    return rows


def report(args, K, W, props, nx, p, Nl, Ng, n, bytes_bs8, host8, ev8, measured_peak, cpu, suite, traffic=None,
           traffic_source=None):
    """The one JSON line: BS8 value and roofline, the suite and the CPU baseline.  `traffic` is the BS8
    kernel's HBM bytes per launch from counters measured in this run; without it the committed
    profile's figure is used when it was made from the same kernel sources, else null."""
    value = bytes_bs8 / host8 / 1e9
    achieved = bytes_bs8 / ev8 / 1e9
    if traffic is None:
        traffic, traffic_source = read_traffic()
    workload = f"BS8 gather+scatter FP64 (Add, Sym, K=1) on a {nx}^3 hex box at p={p} ({Nl} local / {Ng} global nodes)"
    if n is not None:
        workload += f"; suite BS1-BS5 at N=2^{args.log2n}"
    return {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(host8 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: U[-1,1) from the reference LCG (seed 1); lexicographic GLL ids",
        "config": {"workload": workload,
                   "nx": nx, "degree": p, "local_nodes": Nl, "global_nodes": Ng, "vector_n": n,
                   "bytes_per_step": int(bytes_bs8), "GDOFs": round(Ng / host8 / 1e9, 3),
                   "pct_spec_peak": round(100 * value / SPEC_HBM_GBS, 1),
                   "pct_measured_peak": round(100 * achieved / measured_peak, 1) if measured_peak else None,
                   "measured_peak_GBs": round(measured_peak, 1) if measured_peak else None,
                   "traffic_source": traffic_source, "source_sha16": source_hash(),
                   "device": props["name"] or f"{props['arch']} ({props['compute_units']} CUs)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": SPEC_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / SPEC_HBM_GBS, 4),
                     "traffic": round(traffic) if traffic is not None else None},
        "cpu_baseline": cpu,
        "suite": suite,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=64, help="hex elements per side for BS6-BS8")
    ap.add_argument("--degree", type=int, default=7)
    ap.add_argument("--log2n", type=int, default=27, help="BS1-BS5 vector length 2^log2n")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-mesh batched comparison")
    ap.add_argument("--no-degree-sweep", action="store_true", help="skip the p=1..15 sweep at ~1e8 nodes")
    ap.add_argument("--only-batched", action="store_true",
                    help="run the 64-mesh batched comparison alone and print its JSON: the command the committed "
                         "rocprofv3 summary of the batched kernels is taken from")
    ap.add_argument("--only-degree-sweep", action="store_true",
                    help="run the p=1..15 sweep at ~1e8 nodes alone and print its JSON: the command the committed "
                         "rocprofv3 summary of the sweep's kernels is taken from")
    ap.add_argument("--only-bs8", action="store_true",
                    help="time the BS8 step alone (no suite, no CPU baseline): the command the committed "
                         "rocprofv3 summaries of the BS8 kernel are taken from")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure every kernel's HBM traffic")
    ap.add_argument("--no-drivers", action="store_true", help="skip the C++ driver rows (BS0, BS6-BS8 per call)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="internal: the launches the counter passes profile (run by bench.py itself)")
    ap.add_argument("--probe-batched", action="store_true",
                    help="with --traffic-probe: steps of the 64-mesh batch instead of the suite kernels (the "
                         "command of the committed counter summary of the batched kernels)")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args.nx, args.degree, args.log2n, batched=args.probe_batched)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    K, W = args.steps, args.warmup

    import numpy as np

    import streamparanumal_amd as spx
    from streamparanumal_amd import device, ogs, streaming
    from streamparanumal_amd.device import DeviceArray

    ogs.platform_open(0)
    props = device.props()
    log("device", props["name"], props["arch"], props["compute_units"], "CUs")
    if args.only_batched:
        print(json.dumps({"batched_64_meshes": batched_suite(K, W)}), flush=True)
        return
    if args.only_degree_sweep:
        print(json.dumps({"degree_sweep_1e8": degree_sweep(K, W)}), flush=True)
        return

    # ---------------------------------------------------------------- BS6-BS8
    nx, p = args.nx, args.degree
    t0 = time.perf_counter()
    g = ogs.Ogs.hex(nx, nx, nx, p, keep_ids=False)
    Nl, Ng = g.N, g.NgatherGlobal
    log(f"mesh {nx}^3 p={p}: Nl={Nl} Ng={Ng}, setup {time.perf_counter() - t0:.1f}s "
        f"(ogs {g.info['setup_seconds']:.1f}s), blocks {g.info['NblocksT']} x <= {g.info['block_nnz']}")
    q = DeviceArray(Nl)
    device.fill_uniform(q, 1, 0)
    gq = DeviceArray(Ng)

    bytes_bs8 = 4 * (Ng + 1) + 4 * Nl + 16 * Nl
    bytes_bs6 = 4 * (Ng + 1) + 4 * Nl + 8 * Nl + 8 * Ng
    bytes_bs7 = 4 * (Ng + 1) + 4 * Nl + 8 * Ng + 8 * Nl

    host8, ev8 = timed(lambda: g.gather_scatter(q), K, W)
    device.check_kernel_errors()
    log(f"BS8 host {host8 * 1e3:.3f} ms/call, kernel {ev8 * 1e3:.3f} ms -> {bytes_bs8 / ev8 / 1e9:.1f} GB/s")
    suite = {}

    def entry(name, bytes_, dofs, host, ev, extra=None):
        d = {"bytes_per_call": int(bytes_), "host_ms": round(host * 1e3, 4), "kernel_ms": round(ev * 1e3, 4),
             "GBs": round(bytes_ / ev / 1e9, 1), "GBs_host": round(bytes_ / host / 1e9, 1),
             "pct_spec_peak": round(100 * bytes_ / ev / 1e9 / SPEC_HBM_GBS, 1),
             "GDOFs": round(dofs / ev / 1e9, 3)}
        if extra:
            d.update(extra)
        suite[name] = d
        log(name, json.dumps(d))

    entry("BS8_gather_scatter", bytes_bs8, Ng, host8, ev8)
    if args.only_bs8:
        del q, gq, g
        print(json.dumps(report(args, K, W, props, nx, p, Nl, Ng, None, bytes_bs8, host8, ev8, None, None, suite)),
              flush=True)
        return
    h, e = timed(lambda: g.gather(gq, q, 1, spx.ADD, spx.TRANS), K, W)
    entry("BS6_gather", bytes_bs6, Ng, h, e)
    h, e = timed(lambda: g.scatter(q, gq, 1, spx.NOTRANS), K, W)
    entry("BS7_scatter", bytes_bs7, Ng, h, e)
    device.check_kernel_errors()

    # ------------------------------------- BS8 on FP32 / BF16 storage (FP64 sums)
    red = streaming.Reducer(Nl, spx.REDUCE_FAST, 256)

    def rel_err(approx64, exact64):
        nref = (red.norm(exact64), red.value())[1]
        streaming.axpy(1.0, exact64, -1.0, approx64)
        return math.sqrt((red.norm(approx64), red.value())[1] / nref)

    device.fill_uniform(q, 7, 0)
    q64 = DeviceArray(Nl)
    for st, dt, tag, sz in ((spx.STORE_F32, np.float32, "f32", 4), (spx.STORE_BF16, np.uint16, "bf16", 2)):
        qs = DeviceArray(Nl, dt)
        device.convert(q, qs)
        h, e = timed(lambda: g.gather_scatter_mixed(qs), K, W)
        device.check_kernel_errors()
        # error of one application against the FP64 path (bit-exact with the oracle)
        device.convert(q, qs)
        g.gather_scatter_mixed(qs)
        device.convert(qs, q64)
        ref = DeviceArray(Nl)
        streaming.copy(q, ref)
        g.gather_scatter(ref)
        err = rel_err(q64, ref)
        bytes_m = 4 * (Ng + 1) + 4 * Nl + 2 * sz * Nl
        entry(f"BS8_gather_scatter_{tag}", bytes_m, Ng, h, e,
              {"rel_err_vs_f64": float(f"{err:.3e}"), "speedup_vs_f64": round(ev8 / e, 3)})
        del qs, ref
    del q, gq, g, q64, red

    # ---------------------------------------------------------------- BS1-BS5
    n = 1 << args.log2n
    a, b = DeviceArray(n), DeviceArray(n)
    device.fill_uniform(a, 1, 0)
    device.fill_uniform(b, 1, n)
    h, e = timed(lambda: streaming.copy(a, b), K, W)
    entry("BS1_copy", 16 * n, n, h, e)
    # measured peak: the fastest copy variant of this run (streaming or plain accesses, 2-8 pairs per lane)
    variants = {}
    for nt in (1, 0):
        for unroll in (2, 4, 8):
            device.set_tuning("bs1", unroll=unroll, nontemporal=nt)
            _, ev = timed(lambda: streaming.copy(a, b), K, W, settle=0.1)
            variants[f"{'nontemporal' if nt else 'plain'}_u{unroll}"] = round(16 * n / ev / 1e9, 1)
    device.set_tuning("bs1")
    best = max(variants, key=variants.get)
    measured_peak = max(variants[best], 16 * n / e / 1e9)
    suite["BS1_copy"]["variants_GBs"] = variants
    suite["measured_peak"] = {"GBs": round(measured_peak, 1), "variant": best if variants[best] >= 16 * n / e / 1e9
                              else "library default", "pct_spec": round(100 * measured_peak / SPEC_HBM_GBS, 1)}
    log("measured peak", json.dumps(suite["measured_peak"]))
    h, e = timed(lambda: streaming.axpy(0.37, a, -1.21, b), K, W)
    entry("BS2_axpy", 24 * n, n, h, e)
    device.fill_uniform(b, 1, n)  # repeated axpy with |beta| > 1 grows b without bound: fresh inputs
    for mode, tag in ((spx.REDUCE_REFERENCE, "reference"), (spx.REDUCE_FAST, "fast")):
        red = streaming.Reducer(n, mode, 256)
        h, e = timed(lambda: red.norm(a), K, W)
        entry(f"BS3_norm_{tag}", 8 * n, n, h, e)
        h, e = timed(lambda: red.dot(a, b), K, W)
        entry(f"BS4_dot_{tag}", 16 * n, n, h, e)
        h, e = timed(lambda: red.cg_update(a, 1e-9, b), K, W)
        entry(f"BS5_cg_update_{tag}", 24 * n, n, h, e)
    pvec, xvec = DeviceArray(n), DeviceArray(n)
    device.fill_uniform(pvec, 1, 2 * n)
    device.fill_uniform(xvec, 1, 3 * n)
    red = streaming.Reducer(n, spx.REDUCE_FAST, 256)
    h, e = timed(lambda: red.cg_full(1e-9, pvec, a, xvec, b), K, W)
    entry("BS5_cg_full_fast", 48 * n, n, h, e)
    red_ref = streaming.Reducer(n, spx.REDUCE_REFERENCE, 256)
    h, e64 = timed(lambda: red_ref.cg_full(1e-9, pvec, a, xvec, b), K, W)
    entry("BS5_cg_full_reference", 48 * n, n, h, e64)
    # CG update on FP32 / BF16 storage, FP64 rdotr; error of one update vs the FP64 path
    alpha = 0.31
    mix = streaming.MixedCG()
    r64 = DeviceArray(n)
    for st, dt, tag, sz in ((spx.STORE_F32, np.float32, "f32", 4), (spx.STORE_BF16, np.uint16, "bf16", 2)):
        vs = [DeviceArray(n, dt) for _ in range(4)]
        for v, src in zip(vs, (pvec, a, xvec, b)):
            device.convert(src, v)
        h, e = timed(lambda: mix.cg_full(1e-9, *vs), K, W)
        for v, src in zip(vs, (pvec, a, xvec, b)):
            device.convert(src, v)
        mix.cg_full(alpha, *vs)
        rd_mixed = mix.value()
        x0, r0 = DeviceArray(n), DeviceArray(n)
        streaming.copy(xvec, x0)
        streaming.copy(b, r0)
        red_ref.cg_full(alpha, pvec, a, x0, r0)
        rd64 = red_ref.value()
        device.convert(vs[3], r64)
        nref = (red_ref.norm(r0), red_ref.value())[1]
        streaming.axpy(1.0, r0, -1.0, r64)
        err_r = math.sqrt((red_ref.norm(r64), red_ref.value())[1] / nref)
        entry(f"BS5_cg_full_{tag}", 6 * sz * n, n, h, e,
              {"rel_err_r_vs_f64": float(f"{err_r:.3e}"),
               "rel_err_rdotr_vs_f64": float(f"{abs(rd_mixed - rd64) / rd64:.3e}"),
               "speedup_vs_f64_reference": round(e64 / e, 3)})
        del vs, x0, r0
    del pvec, xvec, a, b, r64
    for name, d in suite.items():
        if "GBs" in d and "bytes_per_call" in d:
            d["pct_measured_peak"] = round(100 * d["GBs"] / measured_peak, 1)

    # ---------------------------------------------- C++ drivers: BS0, BS6-BS8 per call
    if not args.no_drivers:
        suite["drivers"] = driver_suite()
        suite["BS0_launch"] = suite["drivers"].pop("BS0_launch")
        log("BS0", json.dumps(suite["BS0_launch"]))

    # C1 output checksum on the GPU (BS8 on 8^3 p=7, seed 0), beside the CPU row's
    g1 = ogs.Ogs.hex(8, 8, 8, 7, keep_ids=False)
    q1 = DeviceArray(g1.N)
    device.fill_uniform(q1, 0, 0)
    g1.gather_scatter(q1)
    c1_checksum = float(np.sum(q1.to_numpy().astype(np.longdouble)))
    del g1, q1

    # ------------------------------------------- precision errors (config-1 and -2 sizes)
    suite["precision_errors"] = precision_errors()

    # ------------------------------------------------------------ degree sweep
    if not args.no_degree_sweep:
        suite["degree_sweep_1e8"] = degree_sweep(K, W)
        suite["element_types_1e8"] = element_types(K, W)

    # ------------------------------------------------------- batched instances
    if not args.no_batched:
        suite["batched_64_meshes"] = batched_suite(K, W)

    # ------------------------------------------------- HBM traffic from counters
    traffic8, traffic_source = None, None
    if not args.no_counters:
        log("counter passes (rocprofv3 FETCH_SIZE, WRITE_SIZE) ...")
        ctr, why = counter_suite(args)
        if ctr is None:
            log("counters:", why)
            suite["counters"] = {"status": why}
        else:
            suite["counters"] = {"status": why, "calibration": ctr.pop("_calibration")}
            for name, t in ctr.items():
                d = suite.get(name)
                if d is None:
                    continue
                d["traffic_bytes"] = int(t["traffic"])
                d["traffic_over_algorithmic"] = round(t["traffic"] / d["bytes_per_call"], 3)
                d["counter_GBs"] = round(t["traffic"] / (d["kernel_ms"] * 1e-3) / 1e9, 1)
                d["counter_pct_spec"] = round(100 * d["counter_GBs"] / SPEC_HBM_GBS, 1)
            if "BS8_gather_scatter" in ctr:
                traffic8, traffic_source = ctr["BS8_gather_scatter"]["traffic"], why
            log("counters", json.dumps({k: suite[k].get("counter_GBs") for k in ctr if k in suite}))

    # ---------------------------------------------------------------- report
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread) ...")
        cpu, cpu_all = cpu_baseline()
        suite["cpu_baseline_all_threads"] = cpu_all
        log("cpu baseline", json.dumps(cpu), json.dumps(cpu_all))
        suite["cpu_rows"] = cpu_rows(c1_checksum)
    out = report(args, K, W, props, nx, p, Nl, Ng, n, bytes_bs8, host8, ev8, measured_peak, cpu, suite, traffic8,
                 traffic_source)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
