"""streamparanumal_amd -- MI355X-native streaming hot path of streamParanumal.

BS0-BS8 (launch latency, copy, axpy, norm, dot, CG update, ogs gather,
scatter, gather+scatter) as hand-written gfx950 HIP kernels behind the C ABI
of libspx.so, with the reference's C++ surface (platform_t, ogs_t, mesh_t,
the BS drivers) in libspxhost.so and bin/BS0..BS8.  This package is the
Python view of the same libraries.
"""
from . import _lib
from ._lib import (ADD, DOUBLE, FLOAT, HALO, INT32, INT64, MAX, MIN, MUL, NOTRANS, REDUCE_FAST,
                   REDUCE_REFERENCE, SIGNED, STORE_BF16, STORE_F32, STORE_F64, SYM, TRANS, UNSIGNED, SpxError,
                   libspx, libspxhost)

__all__ = ["_lib", "libspx", "libspxhost", "SpxError", "ADD", "MUL", "MAX", "MIN", "SYM", "NOTRANS",
           "TRANS", "FLOAT", "DOUBLE", "INT32", "INT64", "UNSIGNED", "SIGNED", "HALO", "REDUCE_REFERENCE",
           "REDUCE_FAST", "STORE_F64", "STORE_F32", "STORE_BF16"]
