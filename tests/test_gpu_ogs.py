"""GPU parity of the ogs gather / scatter / gather+scatter (BS6-BS8) against the oracle.

Bit-exact for every type, op, transpose and field count: each row is folded in
ascending local index on both sides.
"""
import ctypes as C

import numpy as np
import pytest

import streamparanumal_amd as spx
from streamparanumal_amd import device, ogs
from streamparanumal_amd._lib import SpxError, check, libspx
from streamparanumal_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

MESHES = [(1, 1, 1, 1), (2, 1, 1, 1), (2, 2, 2, 2), (3, 2, 1, 3), (5, 3, 2, 4), (4, 4, 4, 15), (8, 8, 8, 7),
          (10, 2, 2, 7)]  # 71 lattice points along x: the 64-point tile blocks of the hex setup split there
DTYPES = [np.float64, np.float32, np.int32, np.int64]
OPS = [spx.ADD, spx.MUL, spx.MAX, spx.MIN]
TRANS = [spx.SYM, spx.NOTRANS, spx.TRANS]


def say(*a):
    print("[case]", *a, flush=True)


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return a.dtype == b.dtype and np.array_equal(a.view(f"u{a.itemsize}"), b.view(f"u{b.itemsize}"))
    return np.array_equal(a, b)


def values(n, dtype, seed, k=1):
    from oracle import pyoracle
    u = pyoracle.uniform(n * k, seed)
    if np.dtype(dtype).kind == "f":
        return u.astype(dtype)
    return np.floor(u * 1000).astype(dtype)


@pytest.fixture(scope="module")
def setups(dev, oracle):
    out = {}
    for m in MESHES:
        ids = ogs.hex_ids(*m)
        out[m] = (ogs.Ogs(ids), oracle.Ogs(ids), ids)
    return out


@pytest.mark.parametrize("mesh", MESHES)
def test_setup_matches_oracle(setups, mesh):
    say("setup", mesh)
    g, o, ids = setups[mesh]
    assert g.NlocalT == o.NrowsT and g.NlocalP == o.NrowsN
    rsT, ciT, _ = g.csr("T")
    rsN, ciN, _ = g.csr("N")
    assert same(rsT, o.rowStartsT) and same(ciT, o.colIdsT)
    assert same(rsN, o.rowStartsN) and same(ciN, o.colIdsN)
    assert same(g.ids, o.ids)


@pytest.mark.parametrize("mesh", MESHES)
def test_hex_setup_equals_generic_on_device(setups, oracle, mesh):
    """ogs_t built in closed form for the box: same CSRs, flags and gather+scatter bits."""
    say("hex setup", mesh)
    g, o, ids = setups[mesh]
    h = ogs.Ogs.hex(*mesh)
    assert h.info["N"] == g.info["N"] and h.NlocalT == g.NlocalT and h.NlocalP == g.NlocalP
    assert same(h.ids, g.ids)
    for which in ("T", "N"):
        for a, b in zip(h.csr(which), g.csr(which)):
            assert same(a, b)
    q = oracle.uniform(g.N, 5)
    dq = DeviceArray.from_numpy(q)
    h.gather_scatter(dq)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), o.gather_scatter(q))


@pytest.mark.parametrize("mesh", MESHES)
def test_bs8_gather_scatter_double(setups, oracle, mesh):
    say("bs8 gs double", mesh)
    g, o, _ = setups[mesh]
    q = oracle.uniform(g.N, 0)
    dq = DeviceArray.from_numpy(q)
    g.gather_scatter(dq)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), o.gather_scatter(q))


@pytest.mark.parametrize("mesh", MESHES)
def test_bs6_gather_bs7_scatter_double(setups, oracle, mesh):
    say("bs6/bs7 double", mesh)
    g, o, _ = setups[mesh]
    q = oracle.uniform(g.N, 0)
    dq, dgq = DeviceArray.from_numpy(q), DeviceArray(g.NlocalT)
    g.gather(dgq, dq, 1, spx.ADD, spx.TRANS)
    gq = dgq.to_numpy()
    assert same(gq, o.gather(q, 1, spx.ADD, spx.TRANS))
    out = DeviceArray(g.N)
    out.fill_zero()
    g.scatter(out, dgq, 1, spx.NOTRANS)
    device.check_kernel_errors()
    assert same(out.to_numpy(), o.scatter(gq, np.zeros(g.N), 1, spx.NOTRANS))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("trans", TRANS)
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_all_types_ops_trans_fields(setups, dtype, op, trans, k):
    mesh = (3, 2, 1, 3)
    say("gs/g/s", mesh, np.dtype(dtype).name, op, trans, k)
    g, o, _ = setups[mesh]
    q = values(g.N, dtype, 5, k)
    if op == spx.MUL:  # keep products finite and exact-ish: values near 1
        q = (np.ones_like(q) + (q % 3 if np.dtype(dtype).kind == "i" else q * 0.01)).astype(dtype)
    dq = DeviceArray.from_numpy(q)
    g.gather_scatter(dq, k, op, trans)
    assert same(dq.to_numpy(), o.gather_scatter(q, k, op, trans))
    rows = g.NlocalP if trans == spx.NOTRANS else g.NlocalT
    dq = DeviceArray.from_numpy(q)
    dgq = DeviceArray(rows * k, dtype)
    g.gather(dgq, dq, k, op, trans)
    gq = dgq.to_numpy()
    assert same(gq, o.gather(q, k, op, trans))
    srows = g.NlocalP if trans == spx.TRANS else g.NlocalT
    gsrc = values(srows, dtype, 6, k)
    base = values(g.N, dtype, 7, k)
    dv, dg = DeviceArray.from_numpy(base), DeviceArray.from_numpy(gsrc)
    g.scatter(dv, dg, k, trans)
    device.check_kernel_errors()
    assert same(dv.to_numpy(), o.scatter(gsrc, base, k, trans))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_gs_nontemporal_values(setups, dtype, op, k):
    """Sym gather+scatter with non-temporal value loads and stores forced on for every node size (one
    pass over K = 1, 2, 4 fields; K = 3 keeps ordinary ones): the oracle's bits."""
    mesh = (3, 2, 1, 3)
    say("gs nontemporal", mesh, np.dtype(dtype).name, op, k)
    g, o, _ = setups[mesh]
    q = op_values(g.N, dtype, 8, k, op)
    g.set_gs_nontemporal(1)
    try:
        assert g.info["gs_nontemporal"] == 0x3E
        dq = DeviceArray.from_numpy(q)
        g.gather_scatter(dq, k, op, spx.SYM)
        device.check_kernel_errors()
        assert same(dq.to_numpy(), o.gather_scatter(q, k, op, spx.SYM))
    finally:
        g.set_gs_nontemporal(-1)
    with pytest.raises(SpxError):
        g.set_gs_nontemporal(2)


@pytest.mark.parametrize("cap", [64, 512, 1024, 3000, 4096])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, "srb"])
def test_result_independent_of_row_blocks(setups, oracle, cap, variant):
    mesh = (5, 3, 2, 4)
    say("block cap", mesh, cap, "variant", variant)
    g, o, _ = setups[mesh]
    q = oracle.uniform(g.N, 11)
    want = o.gather_scatter(q)
    gq_want = o.gather(q, 1, spx.ADD, spx.TRANS)
    g.set_block_nnz(cap)
    g.set_gs_format(variant == "srb")
    device.set_tuning("ogs_gs", variant=-1 if variant == "srb" else variant)
    try:
        dq = DeviceArray.from_numpy(q)
        g.gather_scatter(dq)
        device.check_kernel_errors()
        assert same(dq.to_numpy(), want)
        dq = DeviceArray.from_numpy(q)
        dgq = DeviceArray(g.NlocalT)
        g.gather(dgq, dq, 1, spx.ADD, spx.TRANS)
        assert same(dgq.to_numpy(), gq_want)
        dout = DeviceArray(g.N)
        dout.fill_zero()
        g.scatter(dout, dgq, 1, spx.NOTRANS)
        device.check_kernel_errors()
        assert same(dout.to_numpy(), o.scatter(gq_want, np.zeros(g.N), 1, spx.NOTRANS))
    finally:
        g.set_block_nnz(2048, 2048)
        g.set_gs_format(True)
        device.set_tuning("ogs_gs", variant=-1)


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("variant,cap,xcd", [(0, 1024, 0), (0, 4096, 1), (2, 1024, 1), (2, 4096, 0),
                                             (2, 512, 0), (2, 256, 1), (0, 2048, 32), (2, 1024, 2)])
def test_srb_equals_csr_format(setups, oracle, mesh, dtype, variant, cap, xcd):
    """The sorted-run kernels (256-lane and one-wave tiles) and the T-CSR blocks give the same bits,
    for every op and K = 1, 3."""
    say("srb vs csr", mesh, dtype.__name__, variant, cap, xcd)
    g, o, _ = setups[mesh]
    g.set_block_nnz(2048, gs_cap=cap)
    device.set_tuning("ogs_gs_srb", variant=variant)
    try:
        for k in (1, 3):
            q = values(g.N, dtype, 23, k)
            for op in OPS:
                res = []
                for fmt in (True, False):
                    g.set_gs_format(fmt, xcd)
                    dq = DeviceArray.from_numpy(q)
                    g.gather_scatter(dq, k, op)
                    device.check_kernel_errors()
                    res.append(dq.to_numpy())
                assert same(res[0], res[1])
                assert same(res[0], o.gather_scatter(q, k, op))
    finally:
        g.set_gs_format(True, -1)
        g.set_block_nnz(2048, gs_cap=2048)
        device.set_tuning("ogs_gs_srb", variant=-1)


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("lanes,cap", [(64, 512), (256, 2048), (64, 256)])
def test_srb_gather_scatter_equal_csr(setups, oracle, mesh, dtype, lanes, cap):
    """T-CSR gather (Trans, every op) and scatter (NoTrans) on the sorted-run layout: same bits as
    the CSR kernels and the oracle, K = 1 and 3."""
    say("srb gather/scatter", mesh, dtype.__name__, lanes, cap)
    g, o, _ = setups[mesh]
    g.set_block_nnz(2048, gs_cap=cap)
    g.set_gs_format(True, -1)
    device.set_tuning("ogs_gs_srb", variant=2 if lanes == 64 else 0)
    try:
        for k in (1, 3):
            q = values(g.N, dtype, 29, k)
            for op in OPS:
                outs = []
                for fmt in (True, False):
                    g.set_gather_format(fmt, fmt)
                    dq = DeviceArray.from_numpy(q)
                    dg = DeviceArray(g.NlocalT * k, dtype)
                    g.gather(dg, dq, k, op, spx.TRANS)
                    device.check_kernel_errors()
                    outs.append(dg.to_numpy())
                assert same(outs[0], outs[1])
                assert same(outs[0], o.gather(q, k, op, spx.TRANS))
            gsrc = values(g.NlocalT, dtype, 31, k)
            base = values(g.N, dtype, 32, k)
            outs = []
            for fmt in (True, False):
                g.set_gather_format(fmt, fmt)
                dv = DeviceArray.from_numpy(base)
                g.scatter(dv, DeviceArray.from_numpy(gsrc), k, spx.NOTRANS)
                device.check_kernel_errors()
                outs.append(dv.to_numpy())
            assert same(outs[0], outs[1])
            assert same(outs[0], o.scatter(gsrc, base, k, spx.NOTRANS))
    finally:
        g.set_gather_format(True, True)
        g.set_block_nnz(2048, gs_cap=2048)
        device.set_tuning("ogs_gs_srb", variant=-1)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("cap", [64, 700, 4096])
def test_srb_direct_random_ids(dev, oracle, seed, cap):
    """The C-ABI entry on a layout built from a CSR of random ids (runs of every length)."""
    say("srb direct", seed, cap)
    n = 20000
    ids = (oracle.rand_int(n, 100 + seed) % 6000 + 1).astype(np.int64)
    lone = oracle.rand_int(n, 200 + seed) % 5 == 0
    ids[lone] = 10**6 + np.arange(np.count_nonzero(lone))           # a share of single-member rows
    _, c = ogs.build_csr(ids)
    rs, ci = c["rowStartsT"], c["colIdsT"]
    bs = ogs.block_rows(rs, cap)
    srb = ogs.srb_build(rs, ci, bs, cap)
    q = oracle.uniform(n, seed)
    want = oracle.csr_gs(rs, ci, q)
    for xcd in (-1, -2, 0, 1, 4, 1024):             # XCD remap hints: default (any negative), off, on, blocks
        d = ogs.SrbDevice(srb, xcd_remap=xcd)
        dq = DeviceArray.from_numpy(q)
        d.gather_scatter(dq)
        device.check_kernel_errors()
        assert same(dq.to_numpy(), want)
    for bad in (3, 2048):
        with pytest.raises(SpxError):
            ogs.SrbDevice(srb, xcd_remap=bad).gather_scatter(DeviceArray.from_numpy(q))
    for nt in (0, 0x08, 0x3E):                      # non-temporal values: none, 8-byte nodes, every size
        dq = DeviceArray.from_numpy(q)
        ogs.SrbDevice(srb, nontemporal=nt).gather_scatter(dq)
        device.check_kernel_errors()
        assert same(dq.to_numpy(), want)
    for bad in (1, 0x40):
        with pytest.raises(SpxError):
            ogs.SrbDevice(srb, nontemporal=bad).gather_scatter(DeviceArray.from_numpy(q))
    # gather then scatter on the layout reproduces the gather+scatter
    nrows = rs.size - 1
    dg = DeviceArray(nrows)
    d.gather(dg, DeviceArray.from_numpy(q))
    dq2 = DeviceArray(n)
    dq2.fill_zero()
    d.scatter(dq2, dg)
    device.check_kernel_errors()
    assert same(dq2.to_numpy(), oracle.csr_gs(rs, ci, q))


def op_values(n, dtype, seed, k, op):
    """values(); for MUL near 1, so products over long rows stay finite (and integers do not
    overflow)."""
    q = values(n, dtype, seed, k)
    if op == spx.MUL:
        q = (np.ones_like(q) + (q % 3 if np.dtype(dtype).kind == "i" else q * 0.01)).astype(dtype)
    return q


def signed_ids(n, seed, oracle, ordered=False):
    """Random ids with repeats, negatives and zeros: without unique flagging some groups have several
    positive members and some none.  ordered: the groups with no positive member lie in the second half
    only and every other group's first member is positive, so the rows (positive groups first) ascend
    by first member."""
    ids = (oracle.rand_int(n, 300 + seed) % (n // 3) + 1).astype(np.int64)
    r = oracle.rand_int(n, 400 + seed)
    if ordered:
        h = n // 2
        ids[h:] = -ids[h:]                                         # members of earlier groups, or
        fresh = np.nonzero(r[h:] % 3 == 0)[0] + h
        ids[fresh] = -(n + r[fresh] % (n // 5))                    # groups with no positive member
        ids[r % 11 == 0] = 0
        live = np.nonzero(ids[:h])[0]
        _, first = np.unique(ids[live], return_index=True)        # each group's first member stays positive
        neg = np.zeros(h, bool)
        neg[live] = r[live] % 4 == 0
        neg[live[first]] = False
        ids[:h][neg] *= -1
        return ids
    ids[r % 4 == 0] *= -1
    ids[r % 11 == 0] = 0
    return ids


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("seed,unique,ordered,cap", [(0, False, False, 256), (1, False, False, 4096),
                                                     (2, True, False, 700), (3, False, True, 256),
                                                     (4, False, True, 2048)])
def test_srb_signed_direct(dev, oracle, dtype, op, seed, unique, ordered, cap):
    """The N-CSR forms through the C ABI on a layout with N-member masks: NoTrans and Trans
    gather+scatter, the N-CSR gather and the N-CSR scatter, bit-exact with the oracle's CSR loops."""
    say("srb signed direct", np.dtype(dtype).name, op, seed, unique, ordered, cap)
    n = 20000
    ids = signed_ids(n, seed, oracle, ordered)
    o = oracle.Ogs(ids, kind=spx.SIGNED, unique=unique)
    _, c = ogs.build_csr(ids, spx.SIGNED, unique)
    rs, ci = c["rowStartsT"], c["colIdsT"]
    bs = ogs.block_rows(rs, cap)
    lay = ogs.srb_build(rs, ci, bs, cap, n_csr=(c["rowStartsN"], c["colIdsN"]), ncols=n)
    d = ogs.SrbDevice(lay)
    # without unique flagging, groups with no positive member follow the others: the rows of the tile
    # where they start do not ascend, and only the gather+scatter forms (no row numbers) apply
    assert lay["rows_ascending"] == (unique or ordered)
    if ordered:   # N rows of several members, and T rows with none
        assert c["NrowsN"] < c["NrowsT"] and np.diff(c["rowStartsN"]).max() > 1
    for k in (1, 3):
        q = op_values(n, dtype, seed + 10 * k, k, op)
        for trans in (spx.NOTRANS, spx.TRANS, spx.SYM):
            dq = DeviceArray.from_numpy(q)
            d.gather_scatter_signed(dq, k, op, trans)
            device.check_kernel_errors()
            assert same(dq.to_numpy(), o.gather_scatter(q, k, op, trans)), (k, trans)
        if not lay["rows_ascending"]:
            continue
        dg = DeviceArray(c["NrowsN"] * k, np.dtype(dtype))
        d.gather_signed(dg, DeviceArray.from_numpy(q), k, op)
        device.check_kernel_errors()
        assert same(dg.to_numpy(), o.gather(q, k, op, spx.NOTRANS)), k
        gsrc = values(c["NrowsN"], dtype, seed + 20 * k, k)
        base = values(n, dtype, seed + 30 * k, k)
        dv = DeviceArray.from_numpy(base)
        d.scatter_signed(dv, DeviceArray.from_numpy(gsrc), k)
        device.check_kernel_errors()
        assert same(dv.to_numpy(), o.scatter(gsrc, base, k, spx.TRANS)), k


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("trans", [spx.NOTRANS, spx.TRANS])
@pytest.mark.parametrize("unique", [False, True])
def test_signed_ids_non_sym_through_ogs(dev, oracle, dtype, trans, unique):
    """ogs_t's NoTrans / Trans gather+scatter (and, where the gather is defined, the N-CSR gather and
    scatter) on signed random ids, on the sorted-run layout and on the CSR row blocks."""
    say("signed ogs", np.dtype(dtype).name, trans, unique)
    n = 30000
    ids = signed_ids(n, 7, oracle)
    g, o = ogs.Ogs(ids, kind=spx.SIGNED, unique=unique), oracle.Ogs(ids, kind=spx.SIGNED, unique=unique)
    assert g.gather_defined == unique
    for op in OPS:
        for k in (1, 2):
            q = op_values(n, dtype, 5 + k, k, op)
            want = o.gather_scatter(q, k, op, trans)
            for srb in (True, False):
                g.set_gs_format(srb)
                dq = DeviceArray.from_numpy(q)
                g.gather_scatter(dq, k, op, trans)
                device.check_kernel_errors()
                assert same(dq.to_numpy(), want), (op, k, srb)
            g.set_gs_format(True)
            if not unique:
                continue
            gtrans, strans = (spx.NOTRANS, spx.TRANS)          # the N-CSR forms
            for srb in (True, False):
                g.set_gather_format(srb, srb)
                dg = DeviceArray(g.NlocalP * k, np.dtype(dtype))
                g.gather(dg, DeviceArray.from_numpy(q), k, op, gtrans)
                device.check_kernel_errors()
                assert same(dg.to_numpy(), o.gather(q, k, op, gtrans)), (op, k, srb)
                base = values(n, dtype, 9, k)
                dv = DeviceArray.from_numpy(base)
                g.scatter(dv, dg, k, strans)
                device.check_kernel_errors()
                assert same(dv.to_numpy(), o.scatter(o.gather(q, k, op, gtrans), base, k, strans)), (op, k, srb)
            g.set_gather_format(True, True)


@pytest.mark.parametrize("dtype", [np.float64, np.int32])
@pytest.mark.parametrize("source", ["hex", "signed"])
def test_gs_row_order_does_not_change_results(dev, oracle, dtype, source):
    """The gather+scatter tiles over rows in any order (random keys, the hex box's 3-D blocks, the CSR's
    own order) give the oracle's bits for every op and transpose; the gather keeps the CSR's order."""
    say("gs row order", source, np.dtype(dtype).name)
    if source == "hex":
        g = ogs.Ogs.hex(5, 4, 3, 3)                       # 3-D blocks of grid points by default
        ids = g.ids
        o = oracle.Ogs(ogs.hex_ids(5, 4, 3, 3))
        unique = True
    else:
        ids = signed_ids(20000, 5, oracle, True)
        g, o = ogs.Ogs(ids, kind=spx.SIGNED, unique=False), oracle.Ogs(ids, kind=spx.SIGNED, unique=False)
        unique = False
    keys = oracle.rand_int(g.NlocalT, 77) % 1000
    for order in ("default", "random", "csr"):
        if order == "random":
            g.set_gs_row_keys(keys)
        elif order == "csr":
            g.set_gs_row_keys(None)
        for op in OPS:
            for k in (1, 3):
                q = op_values(g.N, dtype, 40 + k, k, op)
                for trans in TRANS:
                    dq = DeviceArray.from_numpy(q)
                    g.gather_scatter(dq, k, op, trans)
                    device.check_kernel_errors()
                    assert same(dq.to_numpy(), o.gather_scatter(q, k, op, trans)), (order, op, k, trans)
                if unique:
                    dg = DeviceArray(g.NlocalT * k, np.dtype(dtype))
                    g.gather(dg, DeviceArray.from_numpy(q), k, op, spx.TRANS)
                    device.check_kernel_errors()
                    assert same(dg.to_numpy(), o.gather(q, k, op, spx.TRANS)), (order, op, k)
    with pytest.raises(ValueError):
        g.set_gs_row_keys(keys[:-1])


def test_batch_non_sym_acts_part_by_part(dev, oracle):
    """A batch whose first part has groups without positive members: the union keeps each part's
    N-CSR rows at its T-CSR rows (empty N rows fill the gap), so the NoTrans / Trans forms act part by
    part, on the sorted-run layout and on the CSR row blocks."""
    parts_ids = [signed_ids(5000, 11, oracle), signed_ids(4000, 12, oracle)]
    parts = [ogs.Ogs(i, kind=spx.SIGNED, unique=False) for i in parts_ids]
    assert parts[0].NlocalP < parts[0].NlocalT
    b = ogs.Ogs.batch(parts)
    say("batch non-sym", b.N)
    assert b.NlocalT == parts[0].NlocalT + parts[1].NlocalT
    assert b.NlocalP == parts[0].NlocalT + parts[1].NlocalP
    rsN, ciN, _ = b.csr("N")
    rsN0, ciN0, _ = parts[0].csr("N")
    assert np.array_equal(rsN[:parts[0].NlocalP + 1], rsN0)
    assert np.all(rsN[parts[0].NlocalP:parts[0].NlocalT + 1] == rsN0[-1])      # the gap rows are empty
    q = values(b.N, np.float64, 3)
    offs = (0, parts_ids[0].size)
    for trans in (spx.NOTRANS, spx.TRANS):
        want = np.concatenate([oracle.Ogs(i, kind=spx.SIGNED, unique=False).gather_scatter(
            q[off:off + i.size], 1, spx.ADD, trans) for i, off in zip(parts_ids, offs)])
        for srb in (True, False):
            b.set_gs_format(srb)
            dq = DeviceArray.from_numpy(q)
            b.gather_scatter(dq, 1, spx.ADD, trans)
            device.check_kernel_errors()
            assert same(dq.to_numpy(), want), (trans, srb)
    b.set_gs_format(True)


def test_reference_bisection_blocks(setups, oracle):
    """The kernel run on the reference's own row blocks (<=512 entries, ogsOperator.cpp:522-594)."""
    mesh = (8, 8, 8, 7)
    say("reference blocks", mesh)
    g, o, _ = setups[mesh]
    q = oracle.uniform(g.N, 0)
    want = oracle.gs_okl(o.blockStartsT, o.rowStartsT, o.colIdsT, o.rowStartsT, o.colIdsT, q)
    assert same(want, o.gather_scatter(q))
    bs = DeviceArray.from_numpy(o.blockStartsT.astype(np.int32))
    rs = DeviceArray.from_numpy(o.rowStartsT.astype(np.int32))
    ci = DeviceArray.from_numpy(o.colIdsT.astype(np.int32))
    dq = DeviceArray.from_numpy(q)
    check(libspx().spx_ogs_gather_scatter(spx.DOUBLE, spx.ADD, 1, o.blockStartsT.size - 1, bs.ptr, o.NrowsT,
                                          rs.ptr, ci.ptr, o.NrowsT, rs.ptr, ci.ptr, 512, dq.ptr), "gs")
    device.check_kernel_errors()
    assert same(dq.to_numpy(), want)


def test_order_probe_corner_row(dev, oracle):
    """{1e16, 1, -1e16} on an 8-member row: only ascending local order gives 0."""
    ids = ogs.hex_ids(2, 2, 2, 1)
    g, o = ogs.Ogs(ids), oracle.Ogs(ids)
    rs, ci, _ = g.csr("T")
    sizes = np.diff(rs)
    r = int(np.argmax(sizes == 8))
    members = ci[rs[r]:rs[r + 1]]
    q = np.zeros(g.N)
    q[members[0]], q[members[1]], q[members[2]] = 1e16, 1.0, -1e16
    dq = DeviceArray.from_numpy(q)
    g.gather_scatter(dq)
    out = dq.to_numpy()
    assert np.all(out[members] == 0.0)
    assert same(out, o.gather_scatter(q))


def test_oversized_block_is_reported(setups, oracle):
    """A block larger than max_block_nnz is refused by the kernel's guard, not overrun."""
    mesh = (2, 2, 2, 2)
    g, o, _ = setups[mesh]
    say("oversized block", mesh)
    bs = DeviceArray.from_numpy(np.array([0, o.NrowsT], dtype=np.int32))
    rs = DeviceArray.from_numpy(o.rowStartsT.astype(np.int32))
    ci = DeviceArray.from_numpy(o.colIdsT.astype(np.int32))
    q = oracle.uniform(g.N, 1)
    dq = DeviceArray.from_numpy(q)
    check(libspx().spx_ogs_gather_scatter(spx.DOUBLE, spx.ADD, 1, 1, bs.ptr, o.NrowsT, rs.ptr, ci.ptr,
                                          o.NrowsT, rs.ptr, ci.ptr, 64, dq.ptr), "gs")
    with pytest.raises(SpxError):
        device.check_kernel_errors()
    assert same(dq.to_numpy(), q)  # untouched


def test_gather_not_defined_raises(dev):
    ids = np.array([1, 1, 2, 2, 3], dtype=np.int64)  # no unique flagging: two positives per group
    g = ogs.Ogs(ids, kind=spx.SIGNED, unique=False)
    assert not g.gather_defined
    q, gq = DeviceArray.from_numpy(np.ones(5)), DeviceArray(3)
    with pytest.raises(SpxError):
        g.gather(gq, q)


def test_full_size_bs8_64cube_p7(dev, oracle):
    """64^3 elements at p=7 (134M local nodes): CSR invariants, then BS8 bit-exact vs the oracle's
    fold over the same CSR, seed 1."""
    nx, p = 64, 7
    say("full size bs8", nx, p)
    ids = ogs.hex_ids(nx, nx, nx, p)
    g = ogs.Ogs(ids)
    rs, ci, _ = g.csr("T")
    assert g.NlocalT == (nx * p + 1) ** 3 == 90518849
    # invariants fixing the reference's operator: every local node in one row, rows in order of first
    # appearance, members ascending, members share an id
    assert ci.size == ids.size and np.array_equal(np.sort(ci), np.arange(ids.size, dtype=np.int32))
    first = ci[rs[:-1]]
    assert np.all(np.diff(first) > 0)
    row_of = np.repeat(np.arange(g.NlocalT), np.diff(rs))
    assert np.all(np.diff(ci)[np.diff(row_of) == 0] > 0)
    lex = np.abs(g.ids[ci])
    assert np.all(lex[1:][np.diff(row_of) == 0] == lex[:-1][np.diff(row_of) == 0])
    del row_of, lex
    q = DeviceArray(ids.size)
    device.fill_uniform(q, 1, 0)
    g.gather_scatter(q)
    device.check_kernel_errors()
    want = oracle.csr_gs(rs, ci, oracle.uniform(ids.size, 1))
    assert same(q.to_numpy(), want)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("trans", [spx.NOTRANS, spx.TRANS])
@pytest.mark.parametrize("k", [1, 3, 6])
def test_scatter_paths_agree(setups, dtype, trans, k):
    """Scatter through the local->row map and over the CSR row blocks give the oracle's result."""
    mesh = (5, 3, 2, 4)
    say("scatter paths", mesh, np.dtype(dtype).name, trans, k)
    g, o, _ = setups[mesh]
    srows = g.NlocalP if trans == spx.TRANS else g.NlocalT
    gsrc = values(srows, dtype, 8, k)
    base = values(g.N, dtype, 9, k)
    want = o.scatter(gsrc, base, k, trans)
    default = g.info["scatter_by_rows"]
    try:
        for by_rows in (False, True):
            g.set_scatter_path(by_rows)
            dv, dg = DeviceArray.from_numpy(base), DeviceArray.from_numpy(gsrc)
            g.scatter(dv, dg, k, trans)
            device.check_kernel_errors()
            assert same(dv.to_numpy(), want), by_rows
    finally:
        g.set_scatter_path(bool(default))


def test_unflagged_scatter_leaves_zero_ids_alone(dev, oracle):
    """Local nodes with id 0 belong to no row: both scatter paths leave them untouched."""
    ids = ogs.hex_ids(3, 2, 2, 2)
    ids[::7] = 0
    g, o = ogs.Ogs(ids), oracle.Ogs(ids)
    gsrc = oracle.uniform(g.NlocalT, 12)
    base = oracle.uniform(g.N, 13)
    want = o.scatter(gsrc, base, 1, spx.NOTRANS)
    for by_rows in (False, True):
        g.set_scatter_path(by_rows)
        dv, dg = DeviceArray.from_numpy(base), DeviceArray.from_numpy(gsrc)
        g.scatter(dv, dg, 1, spx.NOTRANS)
        assert same(dv.to_numpy(), want)
        assert same(dv.to_numpy()[::7], base[::7])


@pytest.mark.parametrize("p,by_rows", [(1, False), (2, False), (3, False), (7, True), (11, True)])
def test_scatter_path_default_follows_row_length(dev, p, by_rows):
    """Short rows (fewer than 0.63 rows per local node) scatter in local order, long rows by rows."""
    g = ogs.Ogs.hex(3, 3, 3, p, keep_ids=False)
    ratio = g.NlocalT / g.N
    say("scatter default", p, round(ratio, 3))
    assert bool(g.info["scatter_by_rows"]) == by_rows == (ratio >= 0.63)


@pytest.mark.parametrize("mesh", [(2, 2, 1, 2), (3, 2, 2, 3)])
def test_global_to_local_mapping(dev, oracle, mesh):
    """SetupGlobalToLocalMapping (ogsSetup.cpp:833-849): the T-row of every local node, -1 for nodes
    whose id is zero (left out of the operator)."""
    say("global to local", mesh)
    ids = ogs.hex_ids(*mesh).copy()
    ids[::7] = 0
    g = ogs.Ogs(ids)
    o = oracle.Ogs(ids)
    want = np.full(ids.size, -1, dtype=np.int32)
    for r in range(o.NrowsT):
        want[o.colIdsT[o.rowStartsT[r]:o.rowStartsT[r + 1]]] = r
    got = g.global_to_local()
    assert np.array_equal(got, want)
    assert np.all(got[::7] == -1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.int64])
@pytest.mark.parametrize("k", [2, 3, 4])
def test_multi_field_gather_scatter_full_tiles(setups, dtype, k):
    """K fields in one pass (word-major access over whole 64-position chunks) on a mesh of many full
    tiles: bit-exact with the oracle."""
    mesh = (8, 8, 8, 7)
    say("multi-field gs", mesh, np.dtype(dtype).name, k)
    g, o, _ = setups[mesh]
    q = values(g.N, dtype, 12, k)
    dq = DeviceArray.from_numpy(q)
    g.gather_scatter(dq, k)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), o.gather_scatter(q, k))


@pytest.mark.parametrize("tiles", ["setup", "line2", "line1", "csr"])
def test_hex_tile_orders(dev, oracle, tiles):
    """A hex box's gather+scatter under each tile order (the setup's own, line tiles of 2 and 1 node
    rows, the CSR's order), for FP64 and FP32 / BF16 storage: the oracle's bits (x extent of 11
    elements: 78 lattice points, so line tiles are cut along x at cap 1024)."""
    box = (11, 4, 3, 7)
    say("hex tiles", tiles)
    g = ogs.Ogs.hex(*box)
    rs, ci, _ = g.csr("T")
    if tiles != "setup":
        g.set_block_nnz(g.info["block_nnz"], 1024)
    if tiles.startswith("line"):
        grid = np.abs(g.ids[ci[rs[:-1]]]) - 1
        order, cuts = ogs.hex_gs_tiles(*box, int(tiles[-1]), 1024, grid)
        key = np.empty(g.NlocalT, dtype=np.int64)
        key[order] = np.cumsum(cuts) - 1
        g.set_gs_tiles(key)
    elif tiles == "csr":
        g.set_gs_row_keys(None)
    q = oracle.uniform(g.N, 80)
    dq = DeviceArray.from_numpy(q)
    g.gather_scatter(dq)
    device.check_kernel_errors()
    assert same(dq.to_numpy(), oracle.csr_gs(rs, ci, q))
    for st in (spx.STORE_F32, spx.STORE_BF16):
        qs = oracle.to_storage(q, st)
        dq = DeviceArray.from_numpy(qs)
        g.gather_scatter_mixed(dq)
        device.check_kernel_errors()
        assert same(dq.to_numpy(), oracle.gs_mixed(st, rs, ci, qs)), st
