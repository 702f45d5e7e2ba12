"""The allocation-free CG update (spx_bs5_full_noalloc) against spx_bs5_full, through the C ABI.

Two calls in a row per mode, the second on a smaller input, each bit-exact with the allocating entry
point on identical inputs (and, in the reference order, with the oracle): the shared scratch keeps the
first call's partials, which the second must not read."""
import numpy as np
import pytest

import streamparanumal_amd as spx
from streamparanumal_amd import device
from streamparanumal_amd._lib import SpxError, check, libspx
from streamparanumal_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

MODES = [(spx.REDUCE_REFERENCE, 256), (spx.REDUCE_REFERENCE, 512), (spx.REDUCE_REFERENCE, 1024),
         (spx.REDUCE_FAST, 256)]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def one_call(oracle, n, seed, mode, bs, alpha):
    p, ap, x, r = (oracle.uniform(n, seed + i) for i in range(4))
    dp, dap = DeviceArray.from_numpy(p), DeviceArray.from_numpy(ap)
    out = {}
    for which in ("alloc", "noalloc"):
        dx, dr, rd = DeviceArray.from_numpy(x), DeviceArray.from_numpy(r), DeviceArray(1)
        if which == "alloc":
            tmp = DeviceArray(max(1, libspx().spx_reduce_tmp_count(n, mode, bs)))
            check(libspx().spx_bs5_full(n, alpha, dp.ptr, dap.ptr, dx.ptr, dr.ptr, tmp.ptr, rd.ptr, mode, bs),
                  "spx_bs5_full")
        else:
            check(libspx().spx_bs5_full_noalloc(n, alpha, dp.ptr, dap.ptr, dx.ptr, dr.ptr, rd.ptr, mode, bs),
                  "spx_bs5_full_noalloc")
        device.finish()
        device.check_kernel_errors()
        out[which] = (dx.to_numpy(), dr.to_numpy(), rd.to_numpy())
    for a, b in zip(out["alloc"], out["noalloc"]):
        assert np.array_equal(bits(a), bits(b))
    if mode == spx.REDUCE_REFERENCE:
        v, xe, re = oracle.bs5_full(alpha, p, ap, x, r, bs)
        assert np.array_equal(bits(out["noalloc"][0]), bits(xe))
        assert np.array_equal(bits(out["noalloc"][1]), bits(re))
        assert bits(out["noalloc"][2])[0] == bits(np.array([v]))[0]


@pytest.mark.parametrize("mode,bs", MODES)
def test_noalloc_matches_allocating_entry_point(dev, oracle, mode, bs):
    print("[case] noalloc cg update", mode, bs, flush=True)
    one_call(oracle, 3_000_001, 21, mode, bs, 0.37)   # odd length: the scalar tail too
    one_call(oracle, 1_000, 31, mode, bs, -1.21)      # smaller: fewer partials than the call before


def test_noalloc_zero_length_and_refusals(dev):
    rd = DeviceArray(1)
    rd.copy_from(np.array([5.0]))
    check(libspx().spx_bs5_full_noalloc(0, 0.5, None, None, None, None, rd.ptr, spx.REDUCE_REFERENCE, 256),
          "spx_bs5_full_noalloc")
    device.finish()
    assert rd.to_numpy()[0] == 0.0
    v = DeviceArray(8)
    with pytest.raises(SpxError):
        check(libspx().spx_bs5_full_noalloc(8, 0.5, v.ptr, v.ptr, v.ptr, v.ptr, rd.ptr, 7, 256), "bad mode")
    with pytest.raises(SpxError):
        check(libspx().spx_bs5_full_noalloc(8, 0.5, v.ptr, v.ptr, v.ptr, v.ptr, rd.ptr, spx.REDUCE_REFERENCE, 300),
              "bad block size")
