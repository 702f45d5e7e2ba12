"""ctypes bindings of libspx.so (include/spx.h) and libspxhost.so (include/spx_host.h).

The libraries are loaded from ``streamparanumal_amd/lib`` inside this source
tree.  A missing library is an error: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")

SPX_OK = 0
SPX_ERR_ARG, SPX_ERR_HIP, SPX_ERR_NODEVICE, SPX_ERR_UNSUPPORTED, SPX_ERR_KERNEL = -1, -2, -3, -4, -5
FLOAT, DOUBLE, INT32, INT64 = 0, 1, 2, 3
ADD, MUL, MAX, MIN = 0, 1, 2, 3
SYM, NOTRANS, TRANS = 0, 1, 2
UNSIGNED, SIGNED, HALO = 0, 1, 2
REDUCE_REFERENCE, REDUCE_FAST = 0, 1
# storage types of the reduced-precision variants; BF16 vectors are held as
# their raw 16 bits (numpy uint16)
STORE_F64, STORE_F32, STORE_BF16 = 0, 1, 2
MIXED_TMP_COUNT = 2048


def storage_of(dtype):
    """SPX_STORE_* for a numpy dtype: float64, float32, or uint16 (bf16 bits)."""
    import numpy as _np
    d = _np.dtype(dtype)
    if d == _np.float64:
        return STORE_F64
    if d == _np.float32:
        return STORE_F32
    if d == _np.uint16:
        return STORE_BF16
    raise TypeError(f"no storage type for {d}")

vp, i32, i64, u64, dbl, cstr, u8 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_char_p, C.c_uint8
pi32, pi64, pint, pflt = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_float)


class DeviceProps(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("arch", C.c_char * 64), ("device_id", C.c_int),
                ("compute_units", C.c_int), ("clock_khz", C.c_int), ("memory_clock_khz", C.c_int),
                ("memory_bus_width", C.c_int), ("global_mem_bytes", C.c_int64), ("l2_bytes", C.c_int64),
                ("lds_bytes_per_block", C.c_int64), ("warp_size", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("block_threads", C.c_int32), ("xcd_remap", C.c_int32), ("unroll", C.c_int32),
                ("grid_blocks", C.c_int32), ("nontemporal", C.c_int32), ("variant", C.c_int32)]


class OgsInfo(C.Structure):
    _fields_ = [("N", i32), ("Ngather", i32), ("NlocalT", i32), ("NlocalP", i32), ("nnzT", i32), ("nnzN", i32),
                ("NblocksT", i32), ("NblocksN", i32), ("block_nnz", i32), ("gs_block_nnz", i32),
                ("Ngs_blocks", i32), ("gather_defined", i32),
                ("kind", i32), ("unique", i32), ("NgatherGlobal", i64), ("setup_seconds", dbl),
                ("scatter_by_rows", i32), ("mixed_tiles", i32 * 2),
                ("gs_nontemporal", i32)]


class GsTuning(C.Structure):
    _fields_ = [("sorted", i32), ("block_nnz", i32), ("tile_lanes", i32), ("xcd_remap", i32), ("ms", dbl),
                ("candidates", i32)]


class CgTuning(C.Structure):
    _fields_ = [("target_blocks", i32), ("unroll", i32), ("nontemporal", i32), ("ms", dbl), ("candidates", i32)]


class Srb(C.Structure):
    """spx_srb_t: device addresses of a sorted-run tile layout."""
    _fields_ = [("nblocks", i32), ("chunks_per_tile", i32), ("xcd_remap", i32), ("tile_lanes", i32),
                ("nontemporal", i32), ("chunks", vp), ("runIds", vp), ("links", vp)]


# name: (restype, argtypes) -- every function of include/spx.h
SPX_FUNCS = {
    "spx_last_error": (cstr, []),
    "spx_version": (C.c_int, []),
    "spx_device_count": (C.c_int, [pint]),
    "spx_device_open": (C.c_int, [C.c_int]),
    "spx_device_finish": (C.c_int, []),
    "spx_device_props": (C.c_int, [C.POINTER(DeviceProps)]),
    "spx_check_kernel_errors": (C.c_int, []),
    "spx_stream_create": (C.c_int, [C.POINTER(vp)]),
    "spx_stream_destroy": (C.c_int, [vp]),
    "spx_stream_finish": (C.c_int, [vp]),
    "spx_set_stream": (C.c_int, [vp]),
    "spx_get_stream": (vp, []),
    "spx_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "spx_malloc_host": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "spx_free": (C.c_int, [vp]),
    "spx_free_host": (C.c_int, [vp]),
    "spx_memcpy": (C.c_int, [vp, vp, C.c_size_t, C.c_int]),
    "spx_fill_bytes": (C.c_int, [vp, C.c_size_t, u8]),
    "spx_event_create": (C.c_int, [C.POINTER(vp)]),
    "spx_event_destroy": (C.c_int, [vp]),
    "spx_event_record": (C.c_int, [vp]),
    "spx_event_elapsed_ms": (C.c_int, [vp, vp, pflt]),
    "spx_graph_begin": (C.c_int, []),
    "spx_graph_end": (C.c_int, [C.POINTER(vp)]),
    "spx_graph_launch": (C.c_int, [vp]),
    "spx_graph_destroy": (C.c_int, [vp]),
    "spx_fill_uniform_f64": (C.c_int, [i64, vp, u64, u64]),
    "spx_fill_loguniform_f64": (C.c_int, [i64, vp, u64, u64]),
    "spx_convert_f64_to_f32": (C.c_int, [i64, vp, vp]),
    "spx_convert_f64_to_bf16": (C.c_int, [i64, vp, vp]),
    "spx_convert_f32_to_f64": (C.c_int, [i64, vp, vp]),
    "spx_convert_bf16_to_f64": (C.c_int, [i64, vp, vp]),
    "spx_bs0": (C.c_int, [i32, vp, vp]),
    "spx_bs1_copy": (C.c_int, [i64, vp, vp]),
    "spx_bs2_axpy": (C.c_int, [i64, dbl, vp, dbl, vp]),
    "spx_reduce_tmp_count": (i64, [i64, C.c_int, C.c_int]),
    "spx_bs3_norm": (C.c_int, [i64, vp, vp, vp, C.c_int, C.c_int]),
    "spx_bs4_dot": (C.c_int, [i64, vp, vp, vp, vp, C.c_int, C.c_int]),
    "spx_bs5_update": (C.c_int, [i64, vp, dbl, vp, vp, vp, C.c_int, C.c_int]),
    "spx_bs3_1": (C.c_int, [i64, i64, vp, vp, C.c_int]),
    "spx_bs4_1": (C.c_int, [i64, i64, vp, vp, vp, C.c_int]),
    "spx_bs5_1": (C.c_int, [i64, i64, vp, dbl, vp, vp, C.c_int]),
    "spx_reduce_2": (C.c_int, [i64, vp, vp, C.c_int]),
    "spx_bs5_full": (C.c_int, [i64, dbl, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int]),
    "spx_bs5_full_noalloc": (C.c_int, [i64, dbl, vp, vp, vp, vp, vp, C.c_int, C.c_int]),
    "spx_ogs_gather": (C.c_int, [C.c_int, C.c_int, C.c_int, i32, vp, vp, vp, i32, vp, vp]),
    "spx_ogs_scatter": (C.c_int, [C.c_int, C.c_int, i32, vp, vp, vp, i32, vp, vp]),
    "spx_ogs_gather_scatter": (C.c_int, [C.c_int, C.c_int, C.c_int, i32, vp, i32, vp, vp, i32, vp, vp, i32, vp]),
    "spx_ogs_scatter_inverse": (C.c_int, [C.c_int, C.c_int, i32, vp, vp, vp]),
    "spx_ogs_extract": (C.c_int, [C.c_int, C.c_int, i32, vp, vp, vp]),
    "spx_ogs_gather_scatter_srb": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(Srb), vp]),
    "spx_ogs_gather_scatter_srb_mixed": (C.c_int, [C.c_int, C.c_int, C.POINTER(Srb), vp]),
    "spx_ogs_gather_srb": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(Srb), vp, vp]),
    "spx_ogs_scatter_srb": (C.c_int, [C.c_int, C.c_int, C.POINTER(Srb), vp, vp]),
    "spx_ogs_gather_scatter_srb_signed": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Srb), vp, vp]),
    "spx_ogs_gather_srb_signed": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(Srb), vp, C.c_int32, vp, vp]),
    "spx_ogs_scatter_srb_signed": (C.c_int, [C.c_int, C.c_int, C.POINTER(Srb), vp, C.c_int32, vp, vp]),
    "spx_ogs_gather_scatter_mixed": (C.c_int, [C.c_int, C.c_int, i32, vp, vp, vp, i32, vp]),
    "spx_cg_plan_create": (C.c_int, [i32, vp, i32, C.POINTER(vp)]),
    "spx_cg_plan_destroy": (C.c_int, [vp]),
    "spx_cg_plan_blocks": (C.c_int, [vp]),
    "spx_cg_batch": (C.c_int, [C.c_int, vp, C.c_double, vp, vp, vp, vp, vp]),
    "spx_bs5_full_mixed": (C.c_int, [C.c_int, C.c_int64, C.c_double, vp, vp, vp, vp, vp, vp]),
    "spx_set_tuning": (C.c_int, [cstr, C.POINTER(Tuning)]),
    "spx_get_tuning": (C.c_int, [cstr, C.POINTER(Tuning)]),
}

# every function of include/spx_host.h
SPXH_FUNCS = {
    "spxh_last_error": (cstr, []),
    "spxh_platform_open": (C.c_int, [C.c_int]),
    "spxh_ogs_setup": (C.c_int, [i32, vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    "spxh_ogs_destroy": (C.c_int, [vp]),
    "spxh_ogs_info": (C.c_int, [vp, C.POINTER(OgsInfo)]),
    "spxh_ogs_copy_csr": (C.c_int, [vp, C.c_int, vp, vp, vp]),
    "spxh_ogs_set_block_nnz": (C.c_int, [vp, i32]),
    "spxh_ogs_set_gs_block_nnz": (C.c_int, [vp, i32]),
    "spxh_ogs_set_scatter_path": (C.c_int, [vp, C.c_int]),
    "spxh_ogs_global_to_local": (C.c_int, [vp, vp]),
    "spxh_ogs_gather_scatter": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "spxh_ogs_gather_scatter_mixed": (C.c_int, [vp, C.c_int, C.c_int, vp]),
    "spxh_ogs_set_gs_format": (C.c_int, [vp, C.c_int, C.c_int]),
    "spxh_ogs_set_gs_row_keys": (C.c_int, [vp, vp]),
    "spxh_ogs_set_gs_tiles": (C.c_int, [vp, vp]),
    "spxh_ogs_set_gs_mixed_tiles": (C.c_int, [vp, C.c_int, vp]),
    "spxh_ogs_set_gs_nontemporal": (C.c_int, [vp, C.c_int]),
    "spxh_hex_gs_tiles": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i32, vp, vp, vp]),
    "spxh_ogs_setup_batch": (C.c_int, [vp, i32, C.POINTER(vp)]),
    "spxh_ogs_setup_hex": (C.c_int, [i32, i32, i32, C.c_int, C.c_int, C.c_int, vp, C.POINTER(vp)]),
    "spxh_ogs_setup_box": (C.c_int, [C.c_int, i32, i32, i32, C.c_int, C.c_int, C.c_int, vp, C.POINTER(vp)]),
    "spxh_csr_build_hex": (C.c_int, [i32, i32, i32, C.c_int, C.c_int, C.c_int, vp, C.POINTER(vp)]),
    "spxh_ogs_set_gather_format": (C.c_int, [vp, C.c_int, C.c_int]),
    "spxh_ogs_autotune": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(GsTuning)]),
    "spxh_cg_batch_autotune": (C.c_int, [i32, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.POINTER(vp),
                                         C.POINTER(CgTuning)]),
    "spxh_srb_build": (C.c_int, [i32, vp, vp, i32, vp, i32, C.POINTER(vp)]),
    "spxh_srb_sizes": (C.c_int, [vp, vp]),
    "spxh_srb_copy": (C.c_int, [vp, vp, vp, vp, vp]),
    "spxh_srb_destroy": (C.c_int, [vp]),
    "spxh_srb_build_signed": (C.c_int, [i32, vp, vp, i32, vp, i32, i32, i32, vp, vp, C.POINTER(vp)]),
    "spxh_srb_copy_members": (C.c_int, [vp, vp]),
    "spxh_srb_nontemporal": (C.c_int, [vp, C.POINTER(i32)]),
    "spxh_srb_rows_ascending": (C.c_int, [vp, vp]),
    "spxh_ogs_gather": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
    "spxh_ogs_scatter": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp]),
    "spxh_csr_build": (C.c_int, [i32, vp, C.c_int, C.c_int, C.POINTER(vp)]),
    "spxh_csr_sizes": (C.c_int, [vp, vp]),
    "spxh_csr_copy": (C.c_int, [vp, C.c_int, vp, vp]),
    "spxh_csr_destroy": (C.c_int, [vp]),
    "spxh_block_rows": (C.c_int, [i32, vp, i32, vp, vp]),
    "spxh_mesh_hex_ids": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "spxh_mesh_box_shape": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
    "spxh_mesh_box_ids": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
}


class SpxError(RuntimeError):
    """A libspx / libspxhost call returned an error code."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def _bind(lib, table):
    for name, (res, args) in table.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


_spx = None
_spxh = None


def libspx():
    """The loaded libspx.so (raises if it has not been built)."""
    global _spx
    if _spx is None:
        path = os.path.join(LIBDIR, "libspx.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with __graft_entry__.build() "
                              "or `make -C streamparanumal_amd`")
        _spx = C.CDLL(path, mode=C.RTLD_GLOBAL)
        _bind(_spx, SPX_FUNCS)
    return _spx


def libspxhost():
    """The loaded libspxhost.so (raises if it has not been built)."""
    global _spxh
    if _spxh is None:
        libspx()
        path = os.path.join(LIBDIR, "libspxhost.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with __graft_entry__.build()")
        _spxh = C.CDLL(path, mode=C.RTLD_GLOBAL)
        _bind(_spxh, SPXH_FUNCS)
    return _spxh


def check(rc, fn="spx"):
    """Raise SpxError for a nonzero libspx status."""
    if rc != SPX_OK:
        raise SpxError(fn, rc, libspx().spx_last_error().decode(errors="replace"))
    return rc


def checkh(rc, fn="spxh"):
    """Raise SpxError for a nonzero libspxhost status."""
    if rc != SPX_OK:
        raise SpxError(fn, rc, libspxhost().spxh_last_error().decode(errors="replace"))
    return rc
