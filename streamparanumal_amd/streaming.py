"""BS0-BS5 streaming operations on device arrays (BS/BSn/okl/bsN.okl)."""
import numpy as np

from . import _lib
from ._lib import check, checkh, libspx, libspxhost
from .device import DeviceArray


def bs0(n, a, b):
    """bs0.okl:27-39 -- one empty 1x1 launch."""
    check(libspx().spx_bs0(n, a.ptr, b.ptr), "spx_bs0")


def copy(a, b, n=None):
    """BS1: b = a (bs1.okl:27-34)."""
    n = a.n if n is None else n
    check(libspx().spx_bs1_copy(n, a.ptr, b.ptr), "spx_bs1_copy")


def axpy(alpha, x, beta, y, n=None):
    """BS2: y = alpha*x + beta*y, two products and one add (bs2.okl:27-36)."""
    n = x.n if n is None else n
    check(libspx().spx_bs2_axpy(n, alpha, x.ptr, beta, y.ptr), "spx_bs2_axpy")


class Reducer:
    """Scratch for the BS3-BS5 reductions: per-block partials and the result."""

    def __init__(self, n, mode=_lib.REDUCE_REFERENCE, block_size=256):
        self.mode, self.block_size = mode, block_size
        cnt = libspx().spx_reduce_tmp_count(n, mode, block_size)
        if cnt < 0:
            raise ValueError("bad reduction parameters")
        self.tmp = DeviceArray(max(cnt, 1))
        self.out = DeviceArray(1)

    def value(self):
        return float(self.out.to_numpy()[0])

    def norm(self, a, n=None):
        """BS3: sum a^2 (bs3.okl:27-102)."""
        n = a.n if n is None else n
        check(libspx().spx_bs3_norm(n, a.ptr, self.tmp.ptr, self.out.ptr, self.mode, self.block_size),
              "spx_bs3_norm")

    def dot(self, a, b, n=None):
        """BS4: sum a*b (bs4.okl:27-103)."""
        n = a.n if n is None else n
        check(libspx().spx_bs4_dot(n, a.ptr, b.ptr, self.tmp.ptr, self.out.ptr, self.mode, self.block_size),
              "spx_bs4_dot")

    def cg_update(self, Ap, alpha, r, n=None):
        """BS5 kernel form: r += alpha*Ap; rdotr = sum r^2 (bs5.okl:27-106)."""
        n = Ap.n if n is None else n
        check(libspx().spx_bs5_update(n, Ap.ptr, alpha, r.ptr, self.tmp.ptr, self.out.ptr, self.mode,
                                      self.block_size), "spx_bs5_update")

    def cg_full(self, alpha, p, Ap, x, r, n=None):
        """CG update of README.md:49-53: x += alpha p; r -= alpha Ap; rdotr = sum r^2."""
        n = p.n if n is None else n
        check(libspx().spx_bs5_full(n, alpha, p.ptr, Ap.ptr, x.ptr, r.ptr, self.tmp.ptr, self.out.ptr,
                                    self.mode, self.block_size), "spx_bs5_full")


class MixedCG:
    """CG update on FP32 or BF16 (uint16 bits) vectors: x += alpha p and
    r -= alpha Ap in FP32, rdotr = sum of the stored r^2 in FP64.  FP64
    vectors use Reducer.cg_full."""

    def __init__(self):
        self.tmp = DeviceArray(_lib.MIXED_TMP_COUNT)
        self.out = DeviceArray(1)

    def value(self):
        return float(self.out.to_numpy()[0])

    def cg_full(self, alpha, p, Ap, x, r, n=None):
        n = p.n if n is None else n
        st = _lib.storage_of(p.dtype)
        for a in (Ap, x, r):
            if a.dtype != p.dtype:
                raise TypeError("p, Ap, x, r must share one storage type")
        check(libspx().spx_bs5_full_mixed(st, n, alpha, p.ptr, Ap.ptr, x.ptr, r.ptr, self.tmp.ptr,
                                          self.out.ptr), "spx_bs5_full_mixed")


class CgPlan:
    """Batched CG update over segments of back-to-back vectors (spx_cg_batch): x += alpha p,
    r -= alpha Ap per segment, rdotr[s] in FP64.  seg_starts: nseg+1 entry offsets, each a
    multiple of 4.  target_blocks: workgroups launched over the 4096-entry tiles (the default,
    one per tile, was the fastest on a 2.2e9-entry batch); results do not depend on it."""

    def __init__(self, seg_starts, target_blocks=2**31 - 1):
        import ctypes as C
        st = np.ascontiguousarray(seg_starts, dtype=np.int64)
        self.nseg = st.size - 1
        self.h = C.c_void_p()
        check(libspx().spx_cg_plan_create(self.nseg, st.ctypes.data, target_blocks, C.byref(self.h)),
              "spx_cg_plan_create")
        self.nblocks = libspx().spx_cg_plan_blocks(self.h)
        self.rdotr = DeviceArray(self.nseg)

    @classmethod
    def autotuned(cls, seg_starts, p, Ap, x, r, reps=5):
        """Plan and "cg_batch" tuning picked by timing every candidate on these vectors (values kept)."""
        import ctypes as C
        self = cls.__new__(cls)
        st = np.ascontiguousarray(seg_starts, dtype=np.int64)
        self.nseg = st.size - 1
        self.rdotr = DeviceArray(self.nseg)
        self.h = C.c_void_p()
        best = _lib.CgTuning()
        checkh(libspxhost().spxh_cg_batch_autotune(self.nseg, st.ctypes.data, _lib.storage_of(p.dtype), p.ptr, Ap.ptr,
                                                   x.ptr, r.ptr, self.rdotr.ptr, reps, C.byref(self.h),
                                                   C.byref(best)), "spxh_cg_batch_autotune")
        self.nblocks = libspx().spx_cg_plan_blocks(self.h)
        self.tuning = {f: getattr(best, f) for f, _ in _lib.CgTuning._fields_}
        return self

    def update(self, alpha, p, Ap, x, r):
        st = _lib.storage_of(p.dtype)
        check(libspx().spx_cg_batch(st, self.h, alpha, p.ptr, Ap.ptr, x.ptr, r.ptr, self.rdotr.ptr), "spx_cg_batch")

    def values(self):
        return self.rdotr.to_numpy()

    def __del__(self):
        try:
            if self.h.value:
                libspx().spx_cg_plan_destroy(self.h)
        except Exception:
            pass
