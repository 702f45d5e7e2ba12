"""HBM traffic per kernel launch from rocprofv3's FETCH_SIZE and WRITE_SIZE counters.

Each counter runs in a pass of its own (``rocprofv3 --pmc <counter> -- <command>``), the command's
child process does the launches, and the per-dispatch values (KiB) are read back from the CSV the
profiler writes.  The counters are scaled to bytes by what they read for a plain copy of known size
in the same run: a copy of 2^27 doubles reads and writes 1 GiB each (FETCH_SIZE reads half of that
on MI355X, WRITE_SIZE all of it).

    python -m streamparanumal_amd.counters summarize <trace dir> <fetch dir> <write dir>

prints the per-kernel table (durations from a ``--kernel-trace --stats`` run, counter medians and
bytes) that the committed profile summaries quote;

    python -m streamparanumal_amd.counters table <kernel regex> <pmc dir> [<pmc dir> ...]

prints the median per dispatch of every counter of any --pmc passes for the matching kernels.
"""
import csv
import glob
import os
import re
import shutil
import statistics
import subprocess
import sys
import tempfile

CALIBRATION_BYTES = 1 << 30          # a copy of 2^27 doubles: 1 GiB read, 1 GiB written
CALIBRATION_KERNEL = r"k_copy2<4, false>"


def rocprofv3():
    """Path of rocprofv3, or None."""
    path = shutil.which("rocprofv3")
    if path:
        return path
    cand = "/opt/rocm/bin/rocprofv3"
    return cand if os.path.exists(cand) else None


def _rows(pattern_dir, suffix):
    out = []
    for path in sorted(glob.glob(os.path.join(pattern_dir, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            out.extend(csv.DictReader(f))
    return out


def counter_values(outdir):
    """{(kernel name, counter): [value per dispatch in dispatch order]} from a --pmc pass."""
    rows = _rows(outdir, "counter_collection.csv")
    per = {}
    for r in rows:
        name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
        ctr = r.get("Counter_Name") or r.get("Counter-Name") or ""
        val = r.get("Counter_Value") or r.get("Counter-Value") or "0"
        did = int(r.get("Dispatch_Id") or r.get("Dispatch-Id") or 0)
        per.setdefault((name, ctr), []).append((did, float(val)))
    return {k: [v for _, v in sorted(vs)] for k, vs in per.items()}


def kernel_stats(outdir):
    """{kernel name: {"calls", "avg_ns", "min_ns", "max_ns", "pct"}} from a --kernel-trace --stats run."""
    out = {}
    for r in _rows(outdir, "kernel_stats.csv"):
        out[r["Name"]] = {"calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
                          "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]),
                          "pct": float(r["Percentage"])}
    return out


def run_pass(counter, cmd, outdir, timeout=240, env=None):
    """One rocprofv3 --pmc pass over `cmd` (a list; the program itself comes right after --)."""
    prof = rocprofv3()
    if prof is None:
        raise RuntimeError("rocprofv3 not found")
    full = ["timeout", "-s", "KILL", str(timeout), prof, "--pmc", counter, "-f", "csv", "-d", outdir, "-o", "pass",
            "--", *cmd]
    r = subprocess.run(full, capture_output=True, text=True, env=env, cwd=tempfile.gettempdir())
    found = counter_values(outdir)
    if not found:
        raise RuntimeError(f"rocprofv3 --pmc {counter} left no counter rows (rc {r.returncode}): "
                           f"{(r.stderr or r.stdout)[-1500:]}")
    return found


def _median_for(values, counter, pattern):
    """Median of `counter` over the dispatches of kernels whose name matches `pattern`, and their count."""
    rx = re.compile(pattern)
    vals = [v for (name, ctr), vs in values.items() if ctr == counter and rx.search(name) for v in vs]
    if not vals:
        return None, 0
    return statistics.median(vals), len(vals)


def factors(fetch, write):
    """Bytes per KiB of FETCH_SIZE and of WRITE_SIZE, from the plain copy's dispatches."""
    f, _ = _median_for(fetch, "FETCH_SIZE", re.escape(CALIBRATION_KERNEL))
    w, _ = _median_for(write, "WRITE_SIZE", re.escape(CALIBRATION_KERNEL))
    if not f or not w:
        raise RuntimeError("the calibration copy (plain k_copy2 of 2^27 doubles) is missing from the passes")
    return CALIBRATION_BYTES / f, CALIBRATION_BYTES / w, f, w


def traffic(fetch, write, pattern, ff, fw):
    """(read bytes, write bytes, dispatches) per launch of the kernels matching `pattern`."""
    f, nf = _median_for(fetch, "FETCH_SIZE", pattern)
    w, nw = _median_for(write, "WRITE_SIZE", pattern)
    if f is None or w is None:
        return None
    return f * ff, w * fw, min(nf, nw)


def measure(cmd, patterns, workdir=None, timeout=240, env=None):
    """Run the FETCH_SIZE and WRITE_SIZE passes over `cmd` (which must also run the plain calibration
    copy) and return {key: {"read": B, "write": B, "traffic": B, "dispatches": n}} for each
    {key: kernel-name regex} of `patterns`, plus the calibration under "_calibration"."""
    own = workdir is None
    workdir = workdir or tempfile.mkdtemp(prefix="spx_counters_")
    try:
        fetch = run_pass("FETCH_SIZE", cmd, os.path.join(workdir, "fetch"), timeout, env)
        write = run_pass("WRITE_SIZE", cmd, os.path.join(workdir, "write"), timeout, env)
        ff, fw, fk, wk = factors(fetch, write)
        out = {"_calibration": {"fetch_KiB_per_GiB": fk, "write_KiB_per_GiB": wk, "fetch_factor": ff / 1024,
                                "write_factor": fw / 1024}}
        for key, pat in patterns.items():
            t = traffic(fetch, write, pat, ff, fw)
            if t is not None:
                out[key] = {"read": t[0], "write": t[1], "traffic": t[0] + t[1], "dispatches": t[2]}
        return out
    finally:
        if own:
            shutil.rmtree(workdir, ignore_errors=True)


def summarize(trace_dir, fetch_dir, write_dir, top=12):
    """Markdown table: per kernel, traced calls and average duration, counter medians, bytes per launch."""
    stats = kernel_stats(trace_dir) if trace_dir and trace_dir != "-" else {}
    fetch, write = counter_values(fetch_dir), counter_values(write_dir)
    ff, fw, fk, wk = factors(fetch, write)
    names = sorted({n for (n, c) in fetch} | set(stats), key=lambda n: -stats.get(n, {}).get("pct", 0.0))
    lines = [f"calibration: the plain copy of 2^27 doubles reads FETCH_SIZE {fk:.0f} KiB and WRITE_SIZE {wk:.0f} KiB "
             f"for 1 GiB each way, so bytes = FETCH_SIZE x 1024 x {ff / 1024:.3f} and WRITE_SIZE x 1024 x "
             f"{fw / 1024:.3f}", "",
             "| kernel | calls | average us | FETCH_SIZE KiB | WRITE_SIZE KiB | read GB | write GB | counter TB/s |",
             "|---|---|---|---|---|---|---|---|"]
    for n in names[:top]:
        f, _ = _median_for(fetch, "FETCH_SIZE", "^" + re.escape(n) + "$")
        w, _ = _median_for(write, "WRITE_SIZE", "^" + re.escape(n) + "$")
        s = stats.get(n, {})
        avg = s.get("avg_ns")
        rd = f * ff / 1e9 if f is not None else None
        wr = w * fw / 1e9 if w is not None else None
        rate = (rd + wr) / (avg * 1e-9) / 1e3 if (avg and rd is not None and wr is not None) else None
        cell = lambda x, fmt: (fmt.format(x) if x is not None else "-")
        lines.append(f"| `{n[:110]}` | {s.get('calls', '-')} | {cell(avg / 1e3 if avg else None, '{:.1f}')} | "
                     f"{cell(f, '{:.1f}')} | {cell(w, '{:.1f}')} | {cell(rd, '{:.4f}')} | {cell(wr, '{:.4f}')} | "
                     f"{cell(rate, '{:.2f}')} |")
    return "\n".join(lines)


def counter_table(dirs, pattern="."):
    """Markdown table: per kernel matching `pattern`, the median per dispatch of every counter found
    in the --pmc passes under `dirs`."""
    vals = {}
    for d in dirs:
        for (name, ctr), vs in counter_values(d).items():
            if re.search(pattern, name):
                vals.setdefault(name, {})[ctr] = statistics.median(vs)
    ctrs = sorted({c for v in vals.values() for c in v})
    lines = ["| kernel | " + " | ".join(ctrs) + " |", "|---|" + "---|" * len(ctrs)]
    for name, v in sorted(vals.items()):
        lines.append(f"| `{name[:90]}` | " + " | ".join(f"{v[c]:.4g}" if c in v else "-" for c in ctrs) + " |")
    return "\n".join(lines)


if __name__ == "__main__":
    if len(sys.argv) >= 5 and sys.argv[1] == "summarize":
        print(summarize(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]) if len(sys.argv) > 5 else 12))
    elif len(sys.argv) >= 4 and sys.argv[1] == "table":
        print(counter_table(sys.argv[3:], sys.argv[2]))
    else:
        print(__doc__)
        sys.exit(2)
