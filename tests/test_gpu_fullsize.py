"""Full-size GPU parity: the operators the bench times, at the sizes it times them, against the oracle.

* 64^3 hex elements at p=7 (134,217,728 local nodes, 90,518,849 rows): the closed-form operator with
  its lattice-block gather+scatter tiles (the BS8 step), the BS6 Trans gather and BS7 NoTrans scatter
  (both scatter paths), and the gather+scatter on FP32 / BF16 storage.
* Hex boxes at every degree p = 1..15 of about 3e6 local nodes, and the ~1e8-node boxes of the degree
  sweep at p = 1 and p = 15 (elements per side round((1e8/(p+1)^3)^(1/3))).
* The 64-mesh batch of uneven instances (per-side elements U{4..48}, p in {3,5,7,9}, LCG seed 4; 3
  fields, 2,204,149,272 entries): one FP32 step, checked on the largest instance and on every instance
  with entries past 2^31.

Each operator's T-CSR is first pinned against the oracle's lexicographic ids by invariants that fix it
uniquely (every local node in exactly one row, one id per row, one row per id, rows in order of first
appearance, members ascending); the oracle then folds over that CSR.  Bit-exact throughout.
"""
import numpy as np
import pytest

import streamparanumal_amd as spx
from streamparanumal_amd import device, ogs
from streamparanumal_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def pinned_csr(g, oracle, nx, ny, nz, p):
    """g's T-CSR, after checking that it is the reference operator of the hex box's ids."""
    rs, ci, _ = g.csr("T")
    n = nx * ny * nz * (p + 1) ** 3
    ng = (nx * p + 1) * (ny * p + 1) * (nz * p + 1)
    assert g.N == n and g.NlocalT == ng and rs.size == ng + 1 and ci.size == n
    assert rs[0] == 0 and rs[-1] == n
    sizes = np.diff(rs)
    assert np.all(sizes >= 1)
    assert np.all(np.bincount(ci, minlength=n) == 1)       # every local node in exactly one row
    assert np.all(np.diff(ci[rs[:-1]]) > 0)                  # rows in order of first appearance
    row_of = np.repeat(np.arange(ng, dtype=np.int32), sizes)
    inner = row_of[1:] == row_of[:-1]
    del row_of
    assert np.all(np.diff(ci)[inner] > 0)                    # members ascending within a row
    lex = oracle.hex_ids(nx, ny, nz, p)[ci]
    assert np.all(lex[1:][inner] == lex[:-1][inner])         # a row's members share one id
    # ng rows of one id each over ng distinct ids: one row per id
    return rs, ci


# --------------------------------------------------------------- 64^3, p=7
@pytest.fixture(scope="module")
def c2(dev, oracle):
    nx, p = 64, 7
    say("64^3 p=7 operator (closed form, lattice-block tiles)")
    h = ogs.Ogs.hex(nx, nx, nx, p, keep_ids=False)
    rs, ci = pinned_csr(h, oracle, nx, nx, nx, p)
    q0 = oracle.uniform(h.N, 1)
    yield h, rs, ci, q0
    del h


@pytest.mark.timeout(600)
def test_bs8_lattice_tiles_64cube_p7(c2, oracle):
    """The bench's BS8 step: device-filled seed-1 values, one gather+scatter, all 134M nodes."""
    h, rs, ci, q0 = c2
    say("BS8 64^3 p=7 lattice tiles")
    dq = DeviceArray(h.N)
    device.fill_uniform(dq, 1, 0)
    h.gather_scatter(dq)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), oracle.csr_gs(rs, ci, q0))


@pytest.mark.timeout(600)
def test_bs6_gather_64cube_p7(c2, oracle):
    h, rs, ci, q0 = c2
    say("BS6 Trans gather 64^3 p=7")
    dq, dgq = DeviceArray.from_numpy(q0), DeviceArray(h.NlocalT)
    h.gather(dgq, dq, 1, spx.ADD, spx.TRANS)
    device.check_kernel_errors()
    assert same(dgq.to_numpy(), oracle.csr_gather(rs, ci, q0))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("by_rows", [True, False])
def test_bs7_scatter_64cube_p7(c2, oracle, by_rows):
    h, rs, ci, _ = c2
    say("BS7 NoTrans scatter 64^3 p=7, by rows" if by_rows else "BS7 NoTrans scatter 64^3 p=7, local order")
    gv = oracle.uniform(h.NlocalT, 2)
    base = np.full(h.N, -7.0)
    default = h.info["scatter_by_rows"]
    assert default == 1                                      # 0.674 rows per local node: by rows
    try:
        h.set_scatter_path(by_rows)
        dv, dgv = DeviceArray.from_numpy(base), DeviceArray.from_numpy(gv)
        h.scatter(dv, dgv, 1, spx.NOTRANS)
        device.check_kernel_errors()
    finally:
        h.set_scatter_path(bool(default))
    assert same(dv.to_numpy(), oracle.csr_scatter(rs, ci, gv, base))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("inputs", ["uniform", "loguniform"])
@pytest.mark.parametrize("storage", [spx.STORE_F32, spx.STORE_BF16])
def test_bs8_mixed_storage_64cube_p7(c2, oracle, storage, inputs):
    """FP32 / BF16 storage, FP64 row sums rounded once: the bench's mixed BS8 at 64^3 p=7."""
    h, rs, ci, _ = c2
    say("BS8 mixed 64^3 p=7", storage, inputs)
    q64 = (oracle.uniform if inputs == "uniform" else oracle.loguniform)(h.N, 3)
    qs = oracle.to_storage(q64, storage)
    del q64
    dq = DeviceArray.from_numpy(qs)
    h.gather_scatter_mixed(dq)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), oracle.gs_mixed(storage, rs, ci, qs))


# ------------------------------------------------------------ degree sweep
def check_box(oracle, nx, p, seed=2):
    h = ogs.Ogs.hex(nx, nx, nx, p, keep_ids=False)
    rs, ci = pinned_csr(h, oracle, nx, nx, nx, p)
    q0 = oracle.uniform(h.N, seed)
    dq = DeviceArray.from_numpy(q0)
    h.gather_scatter(dq)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), oracle.csr_gs(rs, ci, q0)), "gather+scatter"
    dq.copy_from(q0)
    dgq = DeviceArray(h.NlocalT)
    h.gather(dgq, dq, 1, spx.ADD, spx.TRANS)
    device.check_kernel_errors()
    gq = dgq.to_numpy()
    assert same(gq, oracle.csr_gather(rs, ci, q0)), "gather"
    dq.fill_zero()
    h.scatter(dq, dgq, 1, spx.NOTRANS)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), oracle.csr_scatter(rs, ci, gq, np.zeros(h.N))), "scatter"


@pytest.mark.timeout(300)
@pytest.mark.parametrize("p", list(range(1, 16)))
def test_degree_sweep_3e6(dev, oracle, p):
    nx = int(round((3e6 / (p + 1) ** 3) ** (1.0 / 3.0)))
    say("degree sweep", nx, p)
    check_box(oracle, nx, p)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("p", [1, 15])
def test_degree_sweep_1e8(dev, oracle, p):
    nx = int(round((1e8 / (p + 1) ** 3) ** (1.0 / 3.0)))
    say("degree sweep ~1e8", nx, p)
    check_box(oracle, nx, p)


# ----------------------------------------------------- 64 uneven instances
@pytest.mark.timeout(900)
def test_batch_64_meshes_fp32_step(dev, oracle):
    """One FP32 step of the 64-mesh batch (3 fields): gather+scatter with FP64 row sums, then
    x += alpha p, r -= alpha q and rdotr per instance; the largest instance and every instance with
    entries past 2^31 against the oracle, from the inputs read back before the step."""
    from streamparanumal_amd.batch import Batch, instances
    meshes = instances(64, 4)
    say("batch 64 meshes setup")
    b = Batch(meshes, 3, np.float32)
    assert b.seg_starts[-1] == 2204149272
    sizes = [g.N for g in b.parts]
    check = sorted({int(np.argmax(sizes))} | {i for i in range(len(meshes)) if b.seg_starts[i + 1] > 2 ** 31})
    assert len(check) >= 3 and any(b.seg_starts[i] >= 2 ** 31 for i in check)
    v = b.allocate(seed=5)
    b.make_plan()
    before = {i: {k: v[k].read(int(b.seg_starts[i]), 3 * b.parts[i].N) for k in ("q", "p", "x", "r")}
              for i in check}
    alpha = 0.37
    say("batch step", check)
    b.step(alpha)
    device.finish()
    device.check_kernel_errors()
    rd = b.plan.values()
    for i in check:
        n, p = meshes[i]
        say("batch instance", i, n, p, int(b.seg_starts[i]))
        a, cnt = int(b.seg_starts[i]), 3 * b.parts[i].N
        rs, ci = pinned_csr(b.parts[i], oracle, n, n, n, p)
        inp = before[i]
        q1 = oracle.gs_mixed(spx.STORE_F32, rs, ci, inp["q"], 3)
        assert same(v["q"].read(a, cnt), q1), i
        want_rd, x1, r1 = oracle.cg_mixed(spx.STORE_F32, alpha, inp["p"], q1, inp["x"], inp["r"])
        assert same(v["x"].read(a, cnt), x1), i
        assert same(v["r"].read(a, cnt), r1), i
        # rdotr: FP64 sums of positive terms, each tile's 4096 terms in lanes and a tree, then the
        # segment's tiles in order: recursive summation of depth d <= tiles + 4096 / 256 + 8 is within
        # d * 2^-53 of the exact sum (the oracle's is exact to long-double rounding)
        depth = -(-cnt // 4096) + 16 + 8
        assert abs(rd[i] - want_rd) <= depth * 2.0 ** -53 * abs(want_rd), (i, rd[i], want_rd)
