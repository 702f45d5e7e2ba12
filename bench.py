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
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SPEC_HBM_GBS = 8000.0
METRIC = "HBM GB/s (% peak) + GDOF/s per BS kernel; mixed/batched vs plain FP64"
PROFILE_COUNTERS = os.path.join(ROOT, "profiles", "bs8_gather_scatter_64cube_p7_counters.md")


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
    """HBM bytes per BS8 launch from the committed rocprofv3 counter summary, if present."""
    if not os.path.exists(PROFILE_COUNTERS):
        return None
    text = open(PROFILE_COUNTERS).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


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


def report(args, K, W, props, nx, p, Nl, Ng, n, bytes_bs8, host8, ev8, measured_peak, cpu, suite):
    """The one JSON line: BS8 value and roofline, the suite and the CPU baseline."""
    value = bytes_bs8 / host8 / 1e9
    achieved = bytes_bs8 / ev8 / 1e9
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
                   "device": props["name"] or f"{props['arch']} ({props['compute_units']} CUs)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": SPEC_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / SPEC_HBM_GBS, 4), "traffic": read_traffic()},
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
    args = ap.parse_args()
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
    measured_peak = 16 * n / e / 1e9
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
        d["pct_measured_peak"] = round(100 * d["GBs"] / measured_peak, 1)

    # ---------------------------------------------------------------- BS0
    ta, tb = DeviceArray(1), DeviceArray(1)
    nl = 10000
    for _ in range(100):
        streaming.bs0(1, ta, tb)
    device.finish()
    t0 = time.perf_counter()
    for _ in range(nl):
        streaming.bs0(1, ta, tb)
    device.finish()
    plain = (time.perf_counter() - t0) / nl
    graph = device.Graph.capture(lambda: [streaming.bs0(1, ta, tb) for _ in range(nl)])
    graph.launch()
    device.finish()
    t0 = time.perf_counter()
    graph.launch()
    device.finish()
    replay = (time.perf_counter() - t0) / nl
    suite["BS0_launch"] = {"launches": nl, "plain_us_per_launch": round(plain * 1e6, 3),
                           "graph_us_per_launch": round(replay * 1e6, 3)}
    log("BS0", json.dumps(suite["BS0_launch"]))

    # ------------------------------------------- precision errors (config-1 and -2 sizes)
    suite["precision_errors"] = precision_errors()

    # ------------------------------------------------------------ degree sweep
    if not args.no_degree_sweep:
        suite["degree_sweep_1e8"] = degree_sweep(K, W)
        suite["element_types_1e8"] = element_types(K, W)

    # ------------------------------------------------------- batched instances
    if not args.no_batched:
        suite["batched_64_meshes"] = batched_suite(K, W)

    # ---------------------------------------------------------------- report
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread) ...")
        cpu, cpu_all = cpu_baseline()
        suite["cpu_baseline_all_threads"] = cpu_all
        log("cpu baseline", json.dumps(cpu), json.dumps(cpu_all))
    out = report(args, K, W, props, nx, p, Nl, Ng, n, bytes_bs8, host8, ev8, measured_peak, cpu, suite)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
