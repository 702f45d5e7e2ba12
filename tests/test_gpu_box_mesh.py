"""BS6-BS8 on every element type of the reference's box meshes (-e Tri/Quad/Tet/Hex): the ogs
operator built from the box ids gathers, scatters and gathers+scatters bit-exactly like the oracle,
and the drivers run with -e."""
import os
import subprocess

import numpy as np
import pytest

import streamparanumal_amd as spx
from streamparanumal_amd import device, ogs
from streamparanumal_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

BOXES = [("Tri", 7, 5, 1, 3), ("Tri", 40, 31, 1, 6), ("Quad", 33, 17, 1, 5), ("Tet", 3, 2, 2, 2),
         ("Tet", 9, 8, 7, 4), ("Tet", 5, 4, 6, 9), ("Hex", 6, 5, 4, 5)]
BIN = os.path.join(os.path.dirname(spx.__file__), "bin")


def say(*a):
    print("[case]", *a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.fixture(scope="module")
def setups(dev, oracle):
    out = {}
    for box in BOXES:
        ids = ogs.box_ids(*box)
        out[box] = (ogs.Ogs(ids), oracle.Ogs(ids))
    return out


@pytest.mark.parametrize("box", BOXES)
def test_box_gather_scatter(setups, oracle, box):
    say("box gs", box)
    g, o = setups[box]
    for k in (1, 3):
        q = oracle.uniform(g.N * k, 2)
        dq = DeviceArray.from_numpy(q)
        g.gather_scatter(dq, k)
        device.check_kernel_errors()
        assert same(dq.to_numpy(), o.gather_scatter(q, k))


@pytest.mark.parametrize("box", BOXES)
def test_box_gather_then_scatter(setups, oracle, box):
    say("box gather/scatter", box)
    g, o = setups[box]
    q = oracle.uniform(g.N, 3)
    dq, dgq = DeviceArray.from_numpy(q), DeviceArray(g.NlocalT)
    g.gather(dgq, dq, 1, spx.ADD, spx.TRANS)
    gq = dgq.to_numpy()
    assert same(gq, o.gather(q, 1, spx.ADD, spx.TRANS))
    out = DeviceArray(g.N)
    out.fill_zero()
    g.scatter(out, dgq, 1, spx.NOTRANS)
    device.check_kernel_errors()
    assert same(out.to_numpy(), o.scatter(gq, np.zeros(g.N), 1, spx.NOTRANS))


@pytest.mark.parametrize("bs", ["BS6", "BS7", "BS8"])
@pytest.mark.parametrize("el,dims", [("Tri", ["-nx", "60", "-ny", "50"]), ("Quad", ["-nx", "40", "-ny", "30"]),
                                     ("Tet", ["-nx", "10", "-ny", "9", "-nz", "8"])])
def test_drivers_run_every_element_type(bs, el, dims):
    say("driver", bs, el)
    exe = os.path.join(BIN, bs)
    r = subprocess.run([exe, "-e", el, "-p", "4"] + dims, capture_output=True, text=True, timeout=90)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [ln for ln in r.stdout.splitlines() if ln.startswith(bs + " = [")]
    assert len(line) == 1, r.stdout
    n, dofs = line[0].split("[")[1].split(",")[:2]
    assert int(n) == 4
    nx, ny = int(dims[1]), int(dims[3])
    nz = int(dims[5]) if el == "Tet" else 0
    want = (nx * 4 + 1) * (ny * 4 + 1) * ((nz * 4 + 1) if nz else 1)
    assert int(dofs) == want


@pytest.mark.parametrize("element,nx,ny,nz,p", [("Tri", 30, 21, 1, 2), ("Tri", 17, 13, 1, 3), ("Quad", 40, 25, 1, 1),
                                                ("Quad", 26, 19, 1, 3), ("Quad", 11, 9, 1, 6), ("Tet", 5, 4, 3, 3),
                                                ("Tet", 7, 6, 5, 1), ("Tet", 4, 3, 3, 2), ("Hex", 6, 5, 4, 2)])
def test_mesh_box_setup(dev, oracle, element, nx, ny, nz, p):
    """mesh_t's setup (lattice-block gather+scatter tiles for hex boxes, Tri / Quad up to p = 3, Tet p = 1, 3):
    the operator and flags of the general setup over the box ids, and the oracle's bits for every
    transpose of the gather+scatter, the gather and the scatter."""
    say("mesh box", element, nx, ny, nz, p)
    ids = ogs.box_ids(element, nx, ny, nz, p)
    g, ref, o = ogs.Ogs.box(element, nx, ny, nz, p), ogs.Ogs(ids), oracle.Ogs(ids)
    assert np.array_equal(g.ids, ref.ids)
    for which in ("T", "N"):
        for a, b in zip(g.csr(which), ref.csr(which)):
            assert np.array_equal(a, b)
    for k in (1, 2):
        q = oracle.uniform(g.N * k, 4)
        for trans in (spx.SYM, spx.NOTRANS, spx.TRANS):
            dq = DeviceArray.from_numpy(q)
            g.gather_scatter(dq, k, spx.ADD, trans)
            device.check_kernel_errors()
            assert same(dq.to_numpy(), o.gather_scatter(q, k, spx.ADD, trans)), (k, trans)
        dg = DeviceArray(g.NlocalT * k)
        g.gather(dg, DeviceArray.from_numpy(q), k, spx.ADD, spx.TRANS)
        gq = dg.to_numpy()
        assert same(gq, o.gather(q, k, spx.ADD, spx.TRANS))
        dv = DeviceArray(g.N * k)
        dv.fill_zero()
        g.scatter(dv, dg, k, spx.NOTRANS)
        device.check_kernel_errors()
        assert same(dv.to_numpy(), o.scatter(gq, np.zeros(g.N * k), k, spx.NOTRANS))
