"""GPU parity of the batched uneven-instance path: the block-diagonal union operator, the batched
CG update with per-instance FP64 rdotr, its graph replay and its autotuner, against the oracle
run instance by instance."""
import numpy as np
import pytest

from streamparanumal_amd import STORE_BF16, STORE_F32, STORE_F64, device, ogs
from streamparanumal_amd.batch import Batch
from streamparanumal_amd.device import DeviceArray
from streamparanumal_amd.streaming import CgPlan

pytestmark = pytest.mark.gpu

MESHES = [(2, 3), (1, 5), (3, 3), (2, 7), (1, 9), (4, 3)]
NP = {STORE_F64: np.float64, STORE_F32: np.float32, STORE_BF16: np.uint16}


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def parts(dev):
    return [ogs.Ogs(ogs.hex_ids(n, n, n, p)) for (n, p) in MESHES]


def slices(b):
    return [(int(b.seg_starts[i]), int(b.seg_starts[i + 1])) for i in range(len(b.parts))]


@pytest.mark.parametrize("st", [STORE_F32, STORE_BF16, STORE_F64])
def test_union_gather_scatter_per_instance(parts, oracle, st):
    say("union gs", st)
    b = Batch(MESHES, K=3, dtype=NP[st], parts=parts)
    assert b.union.N == sum(g.N for g in parts)
    q64 = oracle.uniform(3 * b.N, 7)
    q = oracle.to_storage(q64, st) if st != STORE_F64 else q64
    dq = DeviceArray.from_numpy(q)
    b.union.gather_scatter_mixed(dq, 3)
    device.check_kernel_errors()
    got = dq.to_numpy()
    for (a, e), (n, p) in zip(slices(b), MESHES):
        o = oracle.Ogs(ogs.hex_ids(n, n, n, p))
        rs, ci = o.rowStartsT, o.colIdsT
        want = oracle.gs_mixed(st, rs, ci, q[a:e], 3)
        assert np.array_equal(got[a:e].view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("st", [STORE_F32, STORE_F64])
def test_union_of_ordered_parts(dev, oracle, st):
    """Parts from the closed-form hex setup carry their lattice-block tile orders; the union takes
    them side by side, and its gather+scatter still gives each instance's oracle bits."""
    say("union of ordered parts", st)
    b = Batch(MESHES, K=3, dtype=NP[st])                 # parts from Ogs.hex
    q64 = oracle.uniform(3 * b.N, 8)
    q = oracle.to_storage(q64, st) if st != STORE_F64 else q64
    dq = DeviceArray.from_numpy(q)
    b.union.gather_scatter_mixed(dq, 3)
    device.check_kernel_errors()
    got = dq.to_numpy()
    for (a, e), (n, p) in zip(slices(b), MESHES):
        o = oracle.Ogs(ogs.hex_ids(n, n, n, p))
        want = oracle.gs_mixed(st, o.rowStartsT, o.colIdsT, q[a:e], 3)
        assert np.array_equal(got[a:e].view(np.uint8), want.view(np.uint8)), (n, p)
    if st == STORE_F64:
        dq = DeviceArray.from_numpy(q64)
        b.union.gather_scatter(dq, 3)
        device.check_kernel_errors()
        got = dq.to_numpy()
        for (a, e), (n, p) in zip(slices(b), MESHES):
            o = oracle.Ogs(ogs.hex_ids(n, n, n, p))
            assert np.array_equal(got[a:e].view(np.uint64), o.gather_scatter(q64[a:e], 3).view(np.uint64))


@pytest.mark.parametrize("st", [STORE_F32, STORE_BF16, STORE_F64])
@pytest.mark.parametrize("blocks", [1, 7, 64, 1024])
def test_cg_batch_per_segment(dev, oracle, st, blocks):
    say("cg batch", st, blocks)
    lens = [0, 4, 1024, 3, 1 << 16, 4099, 12, 4096, 8196, 4101]
    starts = [0]
    for n in lens[:-1]:
        starts.append(starts[-1] + (n + 3) // 4 * 4)   # segments start on multiples of 4
    starts.append(starts[-1] + lens[-1])               # the last one ends off a multiple of 4
    total = starts[-1]
    vecs64 = [oracle.uniform(total, s) for s in (11, 12, 13, 14)]
    vecs = [oracle.to_storage(v, st) if st != STORE_F64 else v for v in vecs64]
    p, Ap, x, r = (DeviceArray.from_numpy(v) for v in vecs)
    plan = CgPlan(np.array(starts, dtype=np.int64), blocks)
    alpha = 0.3125 + 1e-9
    plan.update(alpha, p, Ap, x, r)
    device.check_kernel_errors()
    xo, ro, rd = x.to_numpy(), r.to_numpy(), plan.values()
    for s in range(len(lens)):
        a, e = starts[s], starts[s + 1]
        if st == STORE_F64:
            xe = vecs[2][a:e] + alpha * vecs[0][a:e]
            re = vecs[3][a:e] - alpha * vecs[1][a:e]
            ve = float(np.sum(re.astype(np.longdouble) ** 2))
        else:
            ve, xe, re = oracle.cg_mixed(st, alpha, vecs[0][a:e], vecs[1][a:e], vecs[2][a:e], vecs[3][a:e])
        assert np.array_equal(xo[a:e], xe) and np.array_equal(ro[a:e], re)
        assert abs(rd[s] - ve) <= 1e-14 * abs(ve)


@pytest.mark.parametrize("st", [STORE_F32, STORE_F64])
def test_cg_batch_bits_do_not_depend_on_grid_or_unroll(dev, oracle, st):
    """Every tile leaves one partial and each segment folds its partials in tile order, so x, r
    and rdotr are the same bits for any grid size, unroll and store kind."""
    say("cg batch grid independence", st)
    from streamparanumal_amd._lib import libspx
    lens = [40000, 4, 123456, 8192]
    starts = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(starts[-1])
    vecs64 = [oracle.uniform(total, s) for s in (51, 52, 53, 54)]
    vecs = [oracle.to_storage(v, st) if st != STORE_F64 else v for v in vecs64]
    results = []
    for blocks in (1, 5, 64, 4096):
        for unroll, nt in ((1, 0), (2, 1), (4, 0)):
            device.set_tuning("cg_batch", unroll=unroll, nontemporal=nt)
            try:
                p, Ap, x, r = (DeviceArray.from_numpy(v) for v in vecs)
                plan = CgPlan(starts, blocks)
                plan.update(0.375, p, Ap, x, r)
                device.check_kernel_errors()
                results.append((x.to_numpy(), r.to_numpy(), plan.values()))
            finally:
                device.set_tuning("cg_batch")  # library defaults again
    for xo, ro, rd in results[1:]:
        assert np.array_equal(xo, results[0][0]) and np.array_equal(ro, results[0][1])
        assert np.array_equal(rd.view(np.uint64), results[0][2].view(np.uint64))


def test_cg_plan_rejects(dev):
    from streamparanumal_amd._lib import SpxError
    with pytest.raises(SpxError):
        CgPlan(np.array([0, 6, 12], dtype=np.int64))     # a segment starting off a multiple of 4
    with pytest.raises(SpxError):
        CgPlan(np.array([0, 8, 4], dtype=np.int64))      # decreasing


def test_graph_replay_matches_eager(parts, oracle):
    say("batch graph replay")
    b = Batch(MESHES, K=3, dtype=np.float32, parts=parts)
    v = b.allocate(seed=21)
    b.make_plan(64)
    init = {k: a.to_numpy() for k, a in v.items()}
    eager = []
    for _ in range(3):
        b.step(0.25)
        device.finish()
        eager.append(({k: a.to_numpy() for k, a in v.items()}, b.plan.values()))
    for k, a in v.items():
        a.copy_from(init[k])
    g = b.capture(0.25)
    for i in range(3):
        g.launch()
        device.finish()
        for k, a in v.items():
            assert np.array_equal(a.to_numpy(), eager[i][0][k]), (i, k)
        assert np.array_equal(b.plan.values(), eager[i][1])


def test_autotune_keeps_results(parts, oracle):
    say("batch autotune")
    b = Batch(MESHES, K=3, dtype=np.float32, parts=parts)
    v = b.allocate(seed=31)
    init = {k: a.to_numpy() for k, a in v.items()}
    b.make_plan(64)
    b.step(0.5)
    ref = {k: a.to_numpy() for k, a in v.items()}
    ref_rd = b.plan.values()
    for k, a in v.items():
        a.copy_from(init[k])
    tuned = b.autotune(reps=2)
    assert tuned["gather_scatter"]["candidates"] > 1 and tuned["cg"]["candidates"] > 1
    for k, a in v.items():      # the autotuner leaves the vectors as they were
        assert np.array_equal(a.to_numpy(), init[k])
    b.step(0.5)
    for k, a in v.items():
        assert np.array_equal(a.to_numpy(), ref[k])
    assert np.array_equal(b.plan.values(), ref_rd)  # rdotr bits do not depend on the plan's grid


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_single_operator_autotune(parts, oracle, dtype):
    say("ogs autotune", dtype.__name__)
    g = parts[3]
    q = oracle.uniform(g.N, 41).astype(dtype)
    want = DeviceArray.from_numpy(q)
    g.gather_scatter(want)
    best = g.autotune(dtype, 1, reps=2)
    assert best["candidates"] >= 10
    got = DeviceArray.from_numpy(q)
    g.gather_scatter(got)
    assert np.array_equal(got.to_numpy(), want.to_numpy())
