"""The BS0-BS8 drivers through their command line: the reference's flags, the long-only additions,
the printed rows, and refusals with a message and exit status 1."""
import os
import subprocess

import pytest

import streamparanumal_amd as spx

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(spx.__file__), "bin")


def run(name, *args):
    print("[case] driver", name, *args, flush=True)
    return subprocess.run([os.path.join(BIN, name), *args], capture_output=True, text=True, timeout=120)


def row(out, name):
    lines = [ln for ln in out.splitlines() if ln.startswith(name + " = [")]
    assert len(lines) == 1, out
    return lines[0]


@pytest.mark.parametrize("name,sc", [("BS1", 16), ("BS2", 24), ("BS3", 8), ("BS4", 16), ("BS5", 24)])
def test_stream_drivers(name, sc):
    r = run(name, "-b", str(sc * (1 << 20)))
    assert r.returncode == 0, r.stderr[-2000:]
    first = row(r.stdout, name).split("[")[1].split()[0]
    assert int(first) == 1 << 20                     # N = bytes / bytes-per-entry


@pytest.mark.parametrize("storage", ["FP64", "FP32", "BF16"])
def test_bs5_full_form_storage(storage):
    r = run("BS5", "--form", "full", "--storage", storage, "--alpha", "0.37", "-b", str(3 * 4 * (1 << 20)))
    assert r.returncode == 0, r.stderr[-2000:]
    dev = row(r.stdout, "BS5device")
    rdotr = float(dev.split("[")[1].split("]")[0].split()[-1])
    assert rdotr > 0.0
    assert (storage + " storage" in dev) or storage == "FP64"


def test_bs5_storage_needs_full_form():
    r = run("BS5", "--storage", "FP32", "-b", "1048576")
    assert r.returncode == 1
    assert "--form full" in r.stdout + r.stderr


@pytest.mark.parametrize("storage", ["FP64", "FP32", "BF16"])
@pytest.mark.parametrize("fields", ["1", "3"])
def test_bs8_storage(storage, fields):
    r = run("BS8", "-nx", "8", "-ny", "8", "-nz", "8", "-p", "7", "--storage", storage, "--fields", fields)
    assert r.returncode == 0, r.stderr[-2000:]
    assert int(row(r.stdout, "BS8").split("[")[1].split(",")[1]) == 57 ** 3
    assert f"{storage} storage" in row(r.stdout, "BS8device")


@pytest.mark.parametrize("name", ["BS6", "BS7"])
def test_gather_scatter_drivers(name):
    r = run(name, "-nx", "10", "-ny", "9", "-nz", "8", "-p", "5")
    assert r.returncode == 0, r.stderr[-2000:]
    assert int(row(r.stdout, name).split("[")[1].split(",")[1]) == 51 * 46 * 41


@pytest.mark.parametrize("name", ["BS6", "BS7", "BS8"])
def test_synced_rows(name):
    """BSnsync: a host finish after every call, as the reference's ogs_t waits inside each call; it
    can only take longer per call than the kernel-only time."""
    r = run(name, "-nx", "8", "-ny", "8", "-nz", "8", "-p", "7", "--settle", "0.05")
    assert r.returncode == 0, r.stderr[-2000:]
    sync = row(r.stdout, name + "sync")
    vals = [float(x) for x in sync.split("[")[1].split("]")[0].split(",")]
    dev = [float(x) for x in row(r.stdout, name + "device").split("[")[1].split("]")[0].split(",")]
    assert int(vals[1]) == 57 ** 3 and vals[2] > 0
    assert vals[2] >= dev[2]
    assert "back-to-back" in sync


def test_bs0_plain_and_graph():
    for extra in ([], ["--graph", "TRUE"]):
        r = run("BS0", "-n", "2000", *extra)
        assert r.returncode == 0, r.stderr[-2000:]
        assert "BS0" in r.stdout


def test_unknown_flag_is_refused():
    r = run("BS1", "--no-such-flag", "1")
    assert r.returncode != 0


def sweep_rows(out, name):
    """Numeric rows of a sweep's MATLAB block `name = [ ... ];`."""
    text = out.split(name + " = [", 1)[1].split("]", 1)[0]
    return [ln.strip().rstrip(";").split() for ln in text.strip().splitlines() if ln.strip()]


def test_byte_sweep():
    mb = 1 << 20
    r = run("BS1", "-bmin", str(16 * mb), "-bmax", str(64 * mb), "-s", str(16 * mb))
    assert r.returncode == 0, r.stderr[-2000:]
    rows = sweep_rows(r.stdout, "BS1")
    assert [int(x[0]) for x in rows] == [mb, 2 * mb, 3 * mb, 4 * mb]   # N = bytes / 16
    assert all(float(x[3]) > 0 for x in rows)


def test_launch_sweep():
    r = run("BS0", "-nmin", "100", "-nmax", "400", "-s", "100")
    assert r.returncode == 0, r.stderr[-2000:]
    rows = sweep_rows(r.stdout, "BS0")
    assert [int(x[0]) for x in rows] == [100, 200, 300, 400]


@pytest.mark.parametrize("name", [f"BS{i}" for i in range(9)])
def test_runbs_invocation(name):
    """BS/runBS.sh runs every driver as `BSn -m HIP` with its default settings."""
    r = run(name, "-m", "HIP")
    assert r.returncode == 0, r.stderr[-2000:]
    assert f"{name} = [" in r.stdout
