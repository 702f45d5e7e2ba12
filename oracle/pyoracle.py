"""ctypes view of liboracle.so (oracle/spx_oracle.h) -- test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

vp, i32, i64, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
pi32 = C.POINTER(C.c_int32)


class OgsStruct(C.Structure):
    _fields_ = [("N", i32), ("NrowsT", i32), ("NrowsN", i32), ("nnzT", i32), ("nnzN", i32),
                ("rowStartsT", pi32), ("colIdsT", pi32), ("rowStartsN", pi32), ("colIdsN", pi32),
                ("NblocksT", i32), ("blockStartsT", pi32), ("NblocksN", i32), ("blockStartsN", pi32),
                ("ids", C.POINTER(C.c_int64)), ("gather_defined", i32), ("error", i32)]


_FUNCS = {
    "oracle_lcg_state": (u64, [u64, u64]),
    "oracle_fill_uniform": (None, [i64, vp, u64, u64]),
    "oracle_fill_loguniform": (None, [i64, vp, u64, u64]),
    "oracle_fill_rand_int": (None, [i64, vp, u64, u64]),
    "oracle_bs1_copy": (None, [i64, vp, vp]),
    "oracle_bs2_axpy": (None, [i64, dbl, vp, dbl, vp]),
    "oracle_bs3_norm": (dbl, [i64, vp, C.c_int, vp]),
    "oracle_bs4_dot": (dbl, [i64, vp, vp, C.c_int, vp]),
    "oracle_bs5_update": (dbl, [i64, vp, dbl, vp, C.c_int, vp]),
    "oracle_bs5_full": (dbl, [i64, dbl, vp, vp, vp, vp, C.c_int]),
    "oracle_sum_longdouble": (dbl, [i64, vp, vp, C.c_int]),
    "oracle_hex_ids_lex": (None, [C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "oracle_hex_ids_minlocal": (None, [C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "oracle_ogs_setup": (C.POINTER(OgsStruct), [i32, vp, C.c_int, C.c_int]),
    "oracle_ogs_free": (None, [C.POINTER(OgsStruct)]),
    "oracle_ogs_gather": (None, [C.POINTER(OgsStruct), C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
    "oracle_ogs_scatter": (None, [C.POINTER(OgsStruct), C.c_int, C.c_int, C.c_int, vp, vp]),
    "oracle_ogs_gather_scatter": (None, [C.POINTER(OgsStruct), C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "oracle_csr_gs": (None, [i32, vp, vp, C.c_int, vp]),
    "oracle_csr_gather": (None, [i32, vp, vp, C.c_int, vp, vp]),
    "oracle_csr_scatter": (None, [i32, vp, vp, C.c_int, vp, vp]),
    "oracle_ogs_gs_okl": (None, [i32, C.c_int, vp, vp, vp, vp, vp, vp]),
    "oracle_bf16_from_double": (C.c_uint16, [dbl]),
    "oracle_bf16_from_float": (C.c_uint16, [C.c_float]),
    "oracle_bf16_to_double": (dbl, [C.c_uint16]),
    "oracle_gs_mixed": (None, [C.c_int, i32, vp, vp, C.c_int, vp]),
    "oracle_cg_mixed": (dbl, [C.c_int, i64, dbl, vp, vp, vp, vp]),
    "oracle_time_bs8": (dbl, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(dbl), C.POINTER(dbl),
                              C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "oracle_time_ogs": (None, [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.POINTER(dbl),
                               C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "oracle_time_stream": (dbl, [C.c_int, i64, C.c_int, C.c_int]),
    "oracle_max_threads": (C.c_int, []),
}

_lib = None


def build():
    """Compile liboracle.so from oracle/spx_oracle.c (gcc, see oracle/Makefile)."""
    subprocess.run(["make", "-C", HERE, "-B", "liboracle.so"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        for name, (res, args) in _FUNCS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


DOUBLE, FLOAT, INT32, INT64 = 1, 0, 2, 3
_TYPES = {np.dtype(np.float32): FLOAT, np.dtype(np.float64): DOUBLE, np.dtype(np.int32): INT32,
          np.dtype(np.int64): INT64}


def lcg_state(seed, skip):
    return lib().oracle_lcg_state(seed, skip)


def uniform(n, seed, skip=0):
    v = np.empty(n, dtype=np.float64)
    lib().oracle_fill_uniform(n, _p(v), seed, skip)
    return v


def loguniform(n, seed, skip=0):
    v = np.empty(n, dtype=np.float64)
    lib().oracle_fill_loguniform(n, _p(v), seed, skip)
    return v


def rand_int(n, seed, skip=0):
    v = np.empty(n, dtype=np.int32)
    lib().oracle_fill_rand_int(n, _p(v), seed, skip)
    return v


def bs1_copy(a):
    b = np.empty_like(a)
    lib().oracle_bs1_copy(a.size, _p(a), _p(b))
    return b


def bs2_axpy(alpha, x, beta, y):
    y = np.array(y, dtype=np.float64)
    lib().oracle_bs2_axpy(x.size, alpha, _p(x), beta, _p(y))
    return y


def bs3_norm(a, block_size=256):
    return lib().oracle_bs3_norm(a.size, _p(a), block_size, None)


def bs4_dot(a, b, block_size=256):
    return lib().oracle_bs4_dot(a.size, _p(a), _p(b), block_size, None)


def bs5_update(Ap, alpha, r, block_size=256):
    r = np.array(r, dtype=np.float64)
    v = lib().oracle_bs5_update(Ap.size, _p(Ap), alpha, _p(r), block_size, None)
    return v, r


def bs5_full(alpha, p, Ap, x, r, block_size=256):
    x = np.array(x, dtype=np.float64)
    r = np.array(r, dtype=np.float64)
    v = lib().oracle_bs5_full(p.size, alpha, _p(p), _p(Ap), _p(x), _p(r), block_size)
    return v, x, r


def sum_longdouble(a, b=None, mode=0):
    return lib().oracle_sum_longdouble(a.size, _p(a), _p(b), mode)


def hex_ids(nx, ny, nz, p, scheme="lex"):
    ids = np.empty(nx * ny * nz * (p + 1) ** 3, dtype=np.int64)
    f = lib().oracle_hex_ids_lex if scheme == "lex" else lib().oracle_hex_ids_minlocal
    f(nx, ny, nz, p, _p(ids))
    return ids


class Ogs:
    """The oracle's ogs setup and host operations (ogsSetup.cpp, ogsOperator.cpp)."""

    def __init__(self, ids, kind=1, unique=True):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        self._o = lib().oracle_ogs_setup(ids.size, _p(ids), kind, 1 if unique else 0)
        o = self._o.contents
        self.error = o.error
        self.N = o.N
        self.NrowsT, self.NrowsN, self.nnzT, self.nnzN = o.NrowsT, o.NrowsN, o.nnzT, o.nnzN
        self.gather_defined = bool(o.gather_defined)
        if self.error:
            return

        def arr(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0, np.int32)

        self.rowStartsT = arr(o.rowStartsT, o.NrowsT + 1)
        self.colIdsT = arr(o.colIdsT, o.nnzT)
        self.rowStartsN = arr(o.rowStartsN, o.NrowsN + 1)
        self.colIdsN = arr(o.colIdsN, o.nnzN)
        self.blockStartsT = arr(o.blockStartsT, o.NblocksT + 1) if o.NblocksT else np.zeros(1, np.int32)
        self.blockStartsN = arr(o.blockStartsN, o.NblocksN + 1) if o.NblocksN else np.zeros(1, np.int32)
        self.ids = np.ctypeslib.as_array(o.ids, shape=(o.N,)).copy() if o.N else np.zeros(0, np.int64)

    def gather(self, v, k=1, op=0, trans=2):
        v = np.ascontiguousarray(v)
        rows = self.NrowsN if trans == 1 else self.NrowsT
        gv = np.zeros(rows * k, dtype=v.dtype)
        lib().oracle_ogs_gather(self._o, _TYPES[v.dtype], op, trans, k, _p(v), _p(gv))
        return gv

    def scatter(self, gv, v, k=1, trans=1):
        v = np.array(v, copy=True)
        gv = np.ascontiguousarray(gv, dtype=v.dtype)
        lib().oracle_ogs_scatter(self._o, _TYPES[v.dtype], trans, k, _p(gv), _p(v))
        return v

    def gather_scatter(self, v, k=1, op=0, trans=0):
        v = np.array(v, copy=True)
        lib().oracle_ogs_gather_scatter(self._o, _TYPES[v.dtype], op, trans, k, _p(v))
        return v

    def __del__(self):
        try:
            lib().oracle_ogs_free(self._o)
        except Exception:
            pass


def gs_okl(blockStarts, gStarts, gIds, sStarts, sIds, q, k=1):
    q = np.array(q, dtype=np.float64)
    bs = np.ascontiguousarray(blockStarts, dtype=np.int32)
    lib().oracle_ogs_gs_okl(bs.size - 1, k, _p(bs), _p(np.ascontiguousarray(gStarts, np.int32)),
                            _p(np.ascontiguousarray(gIds, np.int32)), _p(np.ascontiguousarray(sStarts, np.int32)),
                            _p(np.ascontiguousarray(sIds, np.int32)), _p(q))
    return q


def csr_gs(rowStarts, colIds, q, k=1):
    """Sym Add gather+scatter of q over the given CSR (the host loop of ogsOperator.cpp)."""
    q = np.array(q, dtype=np.float64)
    rs = np.ascontiguousarray(rowStarts, dtype=np.int32)
    ci = np.ascontiguousarray(colIds, dtype=np.int32)
    lib().oracle_csr_gs(rs.size - 1, _p(rs), _p(ci), k, _p(q))
    return q


def csr_gather(rowStarts, colIds, q, k=1):
    """Add gather of q over the given CSR: gv[row] = sum of q over the row's columns in stored order
    (the host loop of ogsOperator.cpp:61-114)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    rs = np.ascontiguousarray(rowStarts, dtype=np.int32)
    ci = np.ascontiguousarray(colIds, dtype=np.int32)
    gv = np.empty((rs.size - 1) * k, dtype=np.float64)
    lib().oracle_csr_gather(rs.size - 1, _p(rs), _p(ci), k, _p(q), _p(gv))
    return gv


def csr_scatter(rowStarts, colIds, gv, q, k=1):
    """q with q[col] = gv[row] for every column of every row (ogsOperator.cpp:223-267)."""
    q = np.array(q, dtype=np.float64)
    gv = np.ascontiguousarray(gv, dtype=np.float64)
    rs = np.ascontiguousarray(rowStarts, dtype=np.int32)
    ci = np.ascontiguousarray(colIds, dtype=np.int32)
    lib().oracle_csr_scatter(rs.size - 1, _p(rs), _p(ci), k, _p(gv), _p(q))
    return q


STORE_F64, STORE_F32, STORE_BF16 = 0, 1, 2


def to_storage(v, storage):
    """Round float64 values to the storage type (float32 array, or bf16 bits as uint16)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    if storage == STORE_F32:
        with np.errstate(over="ignore"):  # past the f32 range: infinity, as the device's conversion
            return v.astype(np.float32)
    if storage == STORE_BF16:
        f = lib().oracle_bf16_from_double
        return np.array([f(float(x)) for x in v], dtype=np.uint16) if v.size < 4096 else _bf16_vec(v)
    return v.copy()


def _bf16_vec(v):
    # exact nearest-even rounding, vectorised (same rule as oracle_bf16_from_double)
    with np.errstate(over="ignore", invalid="ignore"):
        f = v.astype(np.float32)
        lo = f.view(np.uint32) & np.uint32(0xFFFF0000)
        up = np.abs(lo.view(np.float32).astype(np.float64)) > np.abs(v)
        lo = np.where(up, lo - np.uint32(0x10000), lo).astype(np.uint32)
        hi = (lo + np.uint32(0x10000)).astype(np.uint32)
        # above the largest bf16 the next candidate is 2^128 in magnitude (infinity when taken)
        hv = np.where((hi & np.uint32(0x7FFFFFFF)) == np.uint32(0x7F800000), np.copysign(2.0**128, v),
                      hi.view(np.float32).astype(np.float64))
        dl = np.abs(v - lo.view(np.float32).astype(np.float64))
        dh = np.abs(hv - v)
        r = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where((lo >> 16) & 1, hi, lo)))
        return np.where(np.isnan(v), np.uint16(0x7FC0), (r >> 16).astype(np.uint16)).astype(np.uint16)


def from_storage(q, storage):
    if storage == STORE_BF16:
        return (q.astype(np.uint32) << np.uint32(16)).view(np.float32).astype(np.float64)
    return np.asarray(q, dtype=np.float64)


def gs_mixed(storage, rowStarts, colIds, q, k=1):
    q = np.array(q, copy=True)
    rs = np.ascontiguousarray(rowStarts, dtype=np.int32)
    ci = np.ascontiguousarray(colIds, dtype=np.int32)
    lib().oracle_gs_mixed(storage, rs.size - 1, _p(rs), _p(ci), k, _p(q))
    return q


def cg_mixed(storage, alpha, p, Ap, x, r):
    x, r = np.array(x, copy=True), np.array(r, copy=True)
    v = lib().oracle_cg_mixed(storage, p.size, alpha, _p(p), _p(Ap), _p(x), _p(r))
    return v, x, r


def time_bs8(nx, p, reps, threads):
    cs, st = C.c_double(), C.c_double()
    nl, ng = C.c_int64(), C.c_int64()
    t = lib().oracle_time_bs8(nx, p, reps, threads, C.byref(cs), C.byref(st), C.byref(nl), C.byref(ng))
    return {"seconds": t, "checksum": cs.value, "setup_seconds": st.value, "nlocal": nl.value,
            "nglobal": ng.value}


def time_ogs(nx, p, reps, threads):
    """BS6, BS7 and BS8 on one nx^3 hex box at degree p for each thread count in `threads`:
    {"seconds": {threads: [bs6, bs7, bs8] per call}, "checksums": [bs6, bs7, bs8] of the first call's
    output, ...}."""
    thr = np.ascontiguousarray(threads, dtype=np.int32)
    sec = np.zeros(3 * thr.size)
    cs = np.zeros(3)
    st = C.c_double()
    nl, ng = C.c_int64(), C.c_int64()
    lib().oracle_time_ogs(nx, p, reps, thr.size, _p(thr), _p(sec), _p(cs), C.byref(st), C.byref(nl), C.byref(ng))
    return {"seconds": {int(t): sec[3 * i:3 * i + 3].tolist() for i, t in enumerate(thr)},
            "checksums": cs.tolist(), "setup_seconds": st.value, "nlocal": nl.value, "nglobal": ng.value}


def time_stream(which, n, reps, threads):
    return lib().oracle_time_stream(which, n, reps, threads)


def max_threads():
    return lib().oracle_max_threads()
