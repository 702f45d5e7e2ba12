"""GPU parity of the reduced-precision storage variants (FP32 / BF16) against the oracle.

Gather+scatter: FP64 row sums over stored members in ascending local index,
rounded once -- bit-exact against the oracle's restatement.  CG update: FP32
operations without contraction, stored results bit-exact; rdotr (an FP64 sum
in another order than the oracle's long double) within 1e-14 relative.
The error of each variant against the FP64 path is checked to its bound.
"""
import numpy as np
import pytest

from streamparanumal_amd import STORE_BF16, STORE_F32, STORE_F64, device, ogs
from streamparanumal_amd._lib import SpxError, check, libspx
from streamparanumal_amd.device import DeviceArray
from streamparanumal_amd.streaming import MixedCG, Reducer

pytestmark = pytest.mark.gpu

MESHES = [(1, 1, 1, 1), (2, 2, 2, 2), (3, 2, 1, 3), (5, 3, 2, 4), (4, 4, 4, 15), (8, 8, 8, 7)]
STORES = [STORE_F32, STORE_BF16]
NP = {STORE_F32: np.float32, STORE_BF16: np.uint16}


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def setups(dev, oracle):
    out = {}
    for m in MESHES:
        ids = ogs.hex_ids(*m)
        out[m] = (ogs.Ogs(ids), oracle.Ogs(ids))
    return out


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("st", STORES)
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("sorted_runs", [True, False])
def test_gs_mixed_bit_exact(setups, oracle, mesh, st, k, sorted_runs):
    say("gs mixed", mesh, st, k, "sorted-run" if sorted_runs else "csr")
    g, o = setups[mesh]
    q = oracle.to_storage(oracle.uniform(g.N * k, 11), st)
    dq = DeviceArray.from_numpy(q)
    g.set_gs_format(sorted_runs)
    try:
        g.gather_scatter_mixed(dq, k)
    finally:
        g.set_gs_format(True)
    device.check_kernel_errors()
    want = oracle.gs_mixed(st, o.rowStartsT, o.colIdsT, q, k)
    assert np.array_equal(dq.to_numpy().view(np.uint32 if st == STORE_F32 else np.uint16),
                          want.view(np.uint32 if st == STORE_F32 else np.uint16))


@pytest.mark.parametrize("cap", [256, 1024, 2048, 4096])
@pytest.mark.parametrize("st", STORES)
def test_gs_mixed_block_caps(setups, oracle, cap, st):
    say("gs mixed cap", cap, st)
    g, o = setups[(8, 8, 8, 7)]
    g.set_block_nnz(2048, gs_cap=cap)
    try:
        q = oracle.to_storage(oracle.uniform(g.N, 12), st)
        dq = DeviceArray.from_numpy(q)
        g.gather_scatter_mixed(dq)
        device.check_kernel_errors()
        assert np.array_equal(dq.to_numpy(), oracle.gs_mixed(st, o.rowStartsT, o.colIdsT, q))
    finally:
        g.set_block_nnz(2048, gs_cap=2048)


@pytest.mark.parametrize("st", STORES)
def test_gs_mixed_xcd_hints(setups, oracle, st):
    """Every XCD remap hint of the mixed kernel gives the same bits; other values are refused."""
    say("gs mixed xcd", st)
    g, o = setups[(8, 8, 8, 7)]
    q = oracle.to_storage(oracle.uniform(g.N, 14), st)
    want = oracle.gs_mixed(st, o.rowStartsT, o.colIdsT, q)
    try:
        for xcd in (-1, 0, 1, 2, 32, 1024):
            g.set_gs_format(True, xcd)
            dq = DeviceArray.from_numpy(q)
            g.gather_scatter_mixed(dq)
            device.check_kernel_errors()
            assert np.array_equal(dq.to_numpy(), want)
        with pytest.raises(SpxError):
            g.set_gs_format(True, 5)
    finally:
        g.set_gs_format(True, -1)


def test_gs_mixed_f64_is_plain(setups, oracle):
    say("gs mixed f64")
    g, o = setups[(5, 3, 2, 4)]
    q = oracle.uniform(g.N, 13)
    dq = DeviceArray.from_numpy(q)
    g.gather_scatter_mixed(dq)
    assert np.array_equal(dq.to_numpy().view(np.uint64), o.gather_scatter(q).view(np.uint64))


# Hex boxes set up in closed form: FP32 / BF16 tiles of their own at p = 1, 2 (BF16), 7 (both, line
# tiles) and 15 (both, lattice blocks under FP64's line tiles)
HEX_OWN = {(6, 5, 4, 1): [0, 1], (5, 4, 3, 2): [0, 1], (12, 6, 3, 7): [1, 1], (4, 3, 3, 3): [0, 0],
           (3, 2, 2, 15): [1, 1]}


@pytest.fixture(scope="module")
def hex_setups(dev, oracle):
    return {m: (ogs.Ogs.hex(*m), oracle.Ogs(ogs.hex_ids(*m))) for m in HEX_OWN}


def gs_mixed_matches(g, o, oracle, st, k=1, seed=21):
    q = oracle.to_storage(oracle.uniform(g.N * k, seed), st)
    dq = DeviceArray.from_numpy(q)
    g.gather_scatter_mixed(dq, k)
    device.check_kernel_errors()
    want = oracle.gs_mixed(st, o.rowStartsT, o.colIdsT, q, k)
    u = np.uint32 if st == STORE_F32 else np.uint16
    return np.array_equal(dq.to_numpy().view(u), want.view(u))


@pytest.mark.parametrize("mesh", list(HEX_OWN))
@pytest.mark.parametrize("st", STORES)
@pytest.mark.parametrize("k", [1, 3])
def test_gs_mixed_own_tiles(hex_setups, oracle, mesh, st, k):
    """The FP32 / BF16 tiles a hex box's setup gives by degree, bit-exact against the oracle."""
    say("gs mixed own tiles", mesh, st, k)
    g, o = hex_setups[mesh]
    assert g.info["mixed_tiles"] == HEX_OWN[mesh]
    assert gs_mixed_matches(g, o, oracle, st, k)


@pytest.mark.parametrize("mesh", [(12, 6, 3, 7), (3, 2, 2, 15), (4, 3, 3, 3)])
@pytest.mark.parametrize("st", STORES)
@pytest.mark.parametrize("k", [1, 2, 3])
def test_gs_mixed_nontemporal(hex_setups, oracle, mesh, st, k):
    """FP32 / BF16 gather+scatter with non-temporal values forced on (one pass over K = 1, and K = 2
    for FP32; word-major BF16 fields and K = 3 keep ordinary ones) and off: the oracle's bits."""
    say("gs mixed nontemporal", mesh, st, k)
    g, o = hex_setups[mesh]
    try:
        for mode in (1, 0):
            g.set_gs_nontemporal(mode)
            assert gs_mixed_matches(g, o, oracle, st, k, seed=30 + mode)
    finally:
        g.set_gs_nontemporal(-1)


@pytest.mark.parametrize("st", STORES)
def test_gs_mixed_own_tiles_recut(dev, oracle, st):
    """Own tiles survive a capacity change, an FP64 and a mixed autotune, a reset and custom keys."""
    say("gs mixed own tiles recut", st)
    m = (12, 6, 3, 7)
    g, o = ogs.Ogs.hex(*m), oracle.Ogs(ogs.hex_ids(*m))
    g.set_block_nnz(2048, gs_cap=512)
    assert g.info["mixed_tiles"] == [1, 1] and gs_mixed_matches(g, o, oracle, st, seed=22)
    g.set_block_nnz(2048, gs_cap=2048)
    g.autotune(np.float64, reps=1)
    g.autotune(NP[st], reps=1, mixed=True)
    assert g.info["mixed_tiles"] == [1, 1] and gs_mixed_matches(g, o, oracle, st, seed=23)
    g.set_gs_mixed_tiles(st, None)
    assert g.info["mixed_tiles"][st - 1] == 0 and gs_mixed_matches(g, o, oracle, st, seed=24)
    keys = np.random.default_rng(5).integers(0, 97, g.NlocalT)      # arbitrary row groups
    g.set_gs_mixed_tiles(st, keys)
    assert g.info["mixed_tiles"][st - 1] == 1 and gs_mixed_matches(g, o, oracle, st, 2, seed=25)
    with pytest.raises(SpxError):
        g.set_gs_mixed_tiles(STORE_F64, keys)


def inputs(oracle, kind, n, seed):
    """The two input sets of the precision study: U[-1,1), and |x| log-uniform in [1e-6, 1] with random sign."""
    return oracle.uniform(n, seed) if kind == "uniform" else oracle.loguniform(n, seed)


@pytest.mark.parametrize("kind", ["uniform", "loguniform"])
@pytest.mark.parametrize("st,eps", [(STORE_F32, 2.0**-24), (STORE_BF16, 2.0**-8)])
def test_gs_mixed_error_vs_f64(setups, oracle, st, eps, kind):
    say("gs mixed error", st, kind)
    g, o = setups[(8, 8, 8, 7)]
    q = inputs(oracle, kind, g.N, 3)
    exact = o.gather_scatter(q)
    dq = DeviceArray.from_numpy(oracle.to_storage(q, st))
    g.gather_scatter_mixed(dq)
    got = oracle.from_storage(dq.to_numpy(), st)
    bound = eps * o.gather_scatter(np.abs(q)) + eps * np.abs(exact)
    assert np.all(np.abs(got - exact) <= bound * (1 + 1e-12))
    rel = np.linalg.norm(got - exact) / np.linalg.norm(exact)
    print(f"[err] gs storage {st} {kind}: normwise relative error {rel:.3e}", flush=True)
    if st == STORE_F32:
        assert rel <= 1e-6


@pytest.mark.parametrize("n", [1, 3, 4, 5, 1023, 4096, 100003, 1 << 20, 3_000_001])
@pytest.mark.parametrize("st", STORES)
def test_cg_mixed_bit_exact(dev, oracle, n, st):
    say("cg mixed", n, st)
    alpha = 0.4375 + 1e-9
    vecs = [oracle.to_storage(oracle.uniform(n, s), st) for s in (21, 22, 23, 24)]
    p, Ap, x, r = (DeviceArray.from_numpy(v) for v in vecs)
    cg = MixedCG()
    cg.cg_full(alpha, p, Ap, x, r)
    v, xo, ro = oracle.cg_mixed(st, alpha, *vecs)
    assert np.array_equal(x.to_numpy(), xo) and np.array_equal(r.to_numpy(), ro)
    assert abs(cg.value() - v) <= 1e-14 * abs(v)


def test_cg_mixed_empty_and_errors(dev, oracle):
    say("cg mixed edge")
    cg = MixedCG()
    e = DeviceArray(4, np.float32)
    cg.cg_full(1.0, e, e, e, e, n=0)
    assert cg.value() == 0.0
    a = DeviceArray(16, np.float32)
    with pytest.raises(SpxError):  # misaligned: one entry in
        check(libspx().spx_bs5_full_mixed(STORE_F32, 8, 1.0, a.offset(1), a.ptr, a.ptr, a.ptr, cg.tmp.ptr,
                                          cg.out.ptr), "spx_bs5_full_mixed")
    with pytest.raises(SpxError):
        check(libspx().spx_bs5_full_mixed(STORE_F64, 8, 1.0, a.ptr, a.ptr, a.ptr, a.ptr, cg.tmp.ptr,
                                          cg.out.ptr), "spx_bs5_full_mixed")


@pytest.mark.parametrize("kind", ["uniform", "loguniform"])
@pytest.mark.parametrize("st,tol", [(STORE_F32, 1e-6), (STORE_BF16, 2e-2)])
def test_cg_mixed_error_vs_f64(dev, oracle, st, tol, kind):
    say("cg mixed error", st, kind)
    n = 1 << 20
    alpha = 0.31
    vecs64 = [inputs(oracle, kind, n, s) for s in (3, 32, 33, 34)]
    p, Ap, x, r = (DeviceArray.from_numpy(v) for v in vecs64)
    red = Reducer(n)
    red.cg_full(alpha, p, Ap, x, r)
    x64, r64, v64 = x.to_numpy(), r.to_numpy(), red.value()
    ps, Aps, xs, rs = (DeviceArray.from_numpy(oracle.to_storage(v, st)) for v in vecs64)
    cg = MixedCG()
    cg.cg_full(alpha, ps, Aps, xs, rs)
    xm, rm = oracle.from_storage(xs.to_numpy(), st), oracle.from_storage(rs.to_numpy(), st)
    ex = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    er = np.linalg.norm(rm - r64) / np.linalg.norm(r64)
    ed = abs(cg.value() - v64) / v64
    print(f"[err] cg storage {st} {kind}: x {ex:.3e} r {er:.3e} rdotr {ed:.3e}", flush=True)
    assert max(ex, er, ed) <= tol


@pytest.mark.parametrize("st", STORES)
def test_device_conversions(dev, oracle, st):
    say("convert", st)
    n = 200003
    u = oracle.uniform(n, 41) * np.exp2((oracle.rand_int(n, 42) % 40) - 20.0)
    big = (2 - 2.0**-8) * 2.0**127                  # bf16 overflow midpoint (ties to infinity)
    special = [0.0, -0.0, 1.0 + 2.0**-8 + 2.0**-40, 1.0 + 2.0**-8, np.inf, -np.inf, np.nan, -np.nan, big, -big,
               np.nextafter(big, 0), 1e39, -1e300, 3.0e38, 2.0**-133, 2.0**-134, 1.5 * 2.0**-134, 1e-45, -1e-46]
    u[:len(special)] = special
    src = DeviceArray.from_numpy(u)
    dst = DeviceArray(n, NP[st])
    device.convert(src, dst)
    want = oracle.to_storage(u, st)
    bits = np.uint32 if st == STORE_F32 else np.uint16
    assert np.array_equal(dst.to_numpy().view(bits), want.view(bits))
    back = DeviceArray(n)
    device.convert(dst, back)
    assert np.array_equal(back.to_numpy().view(np.uint64), oracle.from_storage(want, st).view(np.uint64))
