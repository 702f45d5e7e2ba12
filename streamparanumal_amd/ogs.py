"""Python mirror of ogs_t (include/ogs.hpp:212-344) over libspxhost.

    ogs = Ogs(ids, kind=SIGNED, unique=True)       # ogs_t::Setup
    ogs.gather_scatter(q, k=1, op=ADD, trans=SYM)  # ogs_t::GatherScatter
    ogs.gather(gq, q, ...); ogs.scatter(q, gq, ...)

Device arrays are streamparanumal_amd.device.DeviceArray; work is enqueued on
the current stream.  The operator lives on the device; host copies of its CSR
are available for inspection.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, checkh, libspx, libspxhost
from .device import ogs_type

_platform_open = False


def platform_open(device_id=0):
    """platform_t in HIP mode on `device_id` (opens the device once)."""
    global _platform_open
    if not _platform_open:
        checkh(libspxhost().spxh_platform_open(device_id), "spxh_platform_open")
        _platform_open = True


def hex_ids(nx, ny, nz, p):
    """Lexicographic GLL-grid ids of a hex box (meshSetupBoxHex3D.cpp:87-130)."""
    n = nx * ny * nz * (p + 1) ** 3
    ids = np.empty(n, dtype=np.int64)
    checkh(libspxhost().spxh_mesh_hex_ids(nx, ny, nz, p, ids.ctypes.data), "spxh_mesh_hex_ids")
    return ids


ELEMENT_TYPES = {"Tri": 3, "Quad": 4, "Tet": 6, "Hex": 12}


def box_shape(element, nx, ny, nz, p):
    """(elements, nodes per element) of a box of `element` ("Tri", "Quad", "Tet", "Hex")."""
    ne = C.c_int64()
    npe = C.c_int()
    checkh(libspxhost().spxh_mesh_box_shape(ELEMENT_TYPES[element], nx, ny, nz, p, C.byref(ne),
                                            C.byref(npe)), "spxh_mesh_box_shape")
    return ne.value, npe.value


def box_ids(element, nx, ny, nz, p):
    """Lattice ids of a box of `element` in the reference's element and node order
    (meshSetupBoxTri2D.cpp, meshSetupBoxQuad2D.cpp, meshSetupBoxTet3D.cpp, meshSetupBoxHex3D.cpp);
    2-D types ignore nz."""
    ne, npe = box_shape(element, nx, ny, nz, p)
    ids = np.empty(ne * npe, dtype=np.int64)
    checkh(libspxhost().spxh_mesh_box_ids(ELEMENT_TYPES[element], nx, ny, nz, p, ids.ctypes.data),
           "spxh_mesh_box_ids")
    return ids


def build_csr(ids, kind=_lib.SIGNED, unique=True):
    """Host half of the setup, no device: returns (ids_after, dict of CSR arrays)."""
    ids = np.ascontiguousarray(ids, dtype=np.int64).copy()
    h = C.c_void_p()
    checkh(libspxhost().spxh_csr_build(ids.size, ids.ctypes.data, kind, 1 if unique else 0, C.byref(h)),
           "spxh_csr_build")
    return ids, _read_csr(h)


def build_csr_hex(nx, ny, nz, p, kind=_lib.SIGNED, unique=True):
    """The same for the hex box's lexicographic ids, in closed form: (ids_after, dict of CSR arrays)."""
    ids = np.empty(nx * ny * nz * (p + 1) ** 3, dtype=np.int64)
    h = C.c_void_p()
    checkh(libspxhost().spxh_csr_build_hex(nx, ny, nz, p, kind, 1 if unique else 0, ids.ctypes.data,
                                           C.byref(h)), "spxh_csr_build_hex")
    return ids, _read_csr(h)


def _read_csr(h):
    try:
        sz = np.zeros(5, dtype=np.int32)
        checkh(libspxhost().spxh_csr_sizes(h, sz.ctypes.data), "spxh_csr_sizes")
        NrowsT, NrowsN, nnzT, nnzN, gd = (int(x) for x in sz)
        out = {"NrowsT": NrowsT, "NrowsN": NrowsN, "gather_defined": bool(gd)}
        for which, tag, rows, nnz in ((0, "T", NrowsT, nnzT), (1, "N", NrowsN, nnzN)):
            rs = np.zeros(rows + 1, dtype=np.int32)
            ci = np.zeros(max(nnz, 1), dtype=np.int32)
            checkh(libspxhost().spxh_csr_copy(h, which, rs.ctypes.data, ci.ctypes.data), "spxh_csr_copy")
            out["rowStarts" + tag] = rs
            out["colIds" + tag] = ci[:nnz]
        return out
    finally:
        libspxhost().spxh_csr_destroy(h)


# spx_srb_chunk_t as a numpy record
SRB_CHUNK = np.dtype([("runMask", "<u8"), ("linkMask", "<u8"), ("runBase", "<i4"), ("linkBase", "<i4"),
                      ("chunkRow", "<i4"), ("nvalid", "<i4")])


def hex_gs_tiles(nx, ny, nz, p, rows_per_line, cap, row_grid):
    """Line-aligned gather+scatter tiles of a hex box (spxh_hex_gs_tiles): (order, cuts) of the rows
    whose lattice points are row_grid."""
    g = np.ascontiguousarray(row_grid, dtype=np.int64)
    order = np.empty(g.size, dtype=np.int32)
    cuts = np.empty(g.size, dtype=np.uint8)
    checkh(libspxhost().spxh_hex_gs_tiles(nx, ny, nz, p, rows_per_line, cap, g.size, g.ctypes.data,
                                          order.ctypes.data, cuts.ctypes.data), "spxh_hex_gs_tiles")
    return order, cuts


def srb_build(row_starts, col_ids, block_starts, cap, n_csr=None, ncols=None):
    """Sorted-run tiles (spx_srb_t) of a CSR and its row blocks, as host arrays.  n_csr = (rowStartsN,
    colIdsN) of a Signed operator whose T-CSR this is (ncols local nodes) adds its N-member masks
    ("memberN")."""
    rs = np.ascontiguousarray(row_starts, dtype=np.int32)
    ci = np.ascontiguousarray(col_ids, dtype=np.int32)
    bs = np.ascontiguousarray(block_starts, dtype=np.int32)
    h = C.c_void_p()
    if n_csr is None:
        checkh(libspxhost().spxh_srb_build(rs.size - 1, rs.ctypes.data, ci.ctypes.data if ci.size else None,
                                           bs.size - 1, bs.ctypes.data, cap, C.byref(h)), "spxh_srb_build")
    else:
        rsn = np.ascontiguousarray(n_csr[0], dtype=np.int32)
        cin = np.ascontiguousarray(n_csr[1], dtype=np.int32)
        checkh(libspxhost().spxh_srb_build_signed(rs.size - 1, rs.ctypes.data, ci.ctypes.data if ci.size else None,
                                                  bs.size - 1, bs.ctypes.data, cap, ncols, rsn.size - 1,
                                                  rsn.ctypes.data, cin.ctypes.data if cin.size else None,
                                                  C.byref(h)), "spxh_srb_build_signed")
    try:
        sz = np.zeros(6, dtype=np.int64)
        checkh(libspxhost().spxh_srb_sizes(h, sz.ctypes.data), "spxh_srb_sizes")
        nb, nnz, nch, nruns, nlinks, cpt = (int(x) for x in sz)
        out = {"chunks": np.zeros(max(nch, 1), SRB_CHUNK), "runIds": np.zeros(max(nruns, 1), np.int32),
               "links": np.zeros(nlinks + 1, np.uint16), "entryStarts": np.zeros(nb + 1, np.int32)}
        checkh(libspxhost().spxh_srb_copy(h, *(out[k].ctypes.data for k in ("chunks", "runIds", "links",
                                                                                "entryStarts"))), "spxh_srb_copy")
        out.update(nblocks=nb, nnz=nnz, nchunks=nch, nruns=nruns, nlinks=nlinks, chunks_per_tile=cpt)
        asc = C.c_int32()
        checkh(libspxhost().spxh_srb_rows_ascending(h, C.byref(asc)), "spxh_srb_rows_ascending")
        out["rows_ascending"] = bool(asc.value)  # the gather and scatter need it
        nt = C.c_int32()
        checkh(libspxhost().spxh_srb_nontemporal(h, C.byref(nt)), "spxh_srb_nontemporal")
        out["nontemporal"] = nt.value  # spx_srb_t.nontemporal as the builder sets it
        if n_csr is not None:
            out["memberN"] = np.zeros(max(nch, 1), np.uint64)
            checkh(libspxhost().spxh_srb_copy_members(h, out["memberN"].ctypes.data), "spxh_srb_copy_members")
            out["nrowsN"] = np.asarray(n_csr[0]).size - 1
        return out
    finally:
        libspxhost().spxh_srb_destroy(h)


class SrbDevice:
    """A sorted-run layout uploaded to the device; `.desc` is the spx_srb_t to pass by reference."""

    def __init__(self, srb, xcd_remap=-1, tile_lanes=0, nontemporal=None):
        from .device import DeviceArray
        self.arrays = {"chunks": DeviceArray.from_numpy(srb["chunks"].view(np.uint8)),
                       "runIds": DeviceArray.from_numpy(srb["runIds"]),
                       "links": DeviceArray.from_numpy(srb["links"])}
        if "memberN" in srb:
            self.arrays["memberN"] = DeviceArray.from_numpy(srb["memberN"])
            self.nrowsN = srb["nrowsN"]
        nt = srb.get("nontemporal", 0) if nontemporal is None else nontemporal
        self.desc = _lib.Srb(srb["nblocks"], srb["chunks_per_tile"], xcd_remap, tile_lanes, nt,
                             *(self.arrays[k].ptr for k in ("chunks", "runIds", "links")))

    def gather_scatter_signed(self, q, k=1, op=_lib.ADD, trans=_lib.NOTRANS):
        check(libspx().spx_ogs_gather_scatter_srb_signed(ogs_type(q.dtype), op, trans, k, C.byref(self.desc),
                                                         self.arrays["memberN"].ptr, q.ptr),
              "spx_ogs_gather_scatter_srb_signed")

    def gather_signed(self, gq, q, k=1, op=_lib.ADD):
        """The N-CSR (NoTrans) gather."""
        check(libspx().spx_ogs_gather_srb_signed(ogs_type(q.dtype), op, k, C.byref(self.desc),
                                                 self.arrays["memberN"].ptr, self.nrowsN, q.ptr, gq.ptr),
              "spx_ogs_gather_srb_signed")

    def scatter_signed(self, q, gq, k=1):
        """The N-CSR (Trans) scatter."""
        check(libspx().spx_ogs_scatter_srb_signed(ogs_type(q.dtype), k, C.byref(self.desc),
                                                  self.arrays["memberN"].ptr, self.nrowsN, gq.ptr, q.ptr),
              "spx_ogs_scatter_srb_signed")

    def gather_scatter(self, q, k=1, op=_lib.ADD):
        check(libspx().spx_ogs_gather_scatter_srb(ogs_type(q.dtype), op, k, C.byref(self.desc), q.ptr),
              "spx_ogs_gather_scatter_srb")

    def gather(self, gq, q, k=1, op=_lib.ADD):
        check(libspx().spx_ogs_gather_srb(ogs_type(q.dtype), op, k, C.byref(self.desc), q.ptr, gq.ptr),
              "spx_ogs_gather_srb")

    def scatter(self, q, gq, k=1):
        check(libspx().spx_ogs_scatter_srb(ogs_type(q.dtype), k, C.byref(self.desc), gq.ptr, q.ptr),
              "spx_ogs_scatter_srb")


def block_rows(row_starts, cap):
    """Greedy row blocks of at most `cap` entries (the product's partition)."""
    rs = np.ascontiguousarray(row_starts, dtype=np.int32)
    nrows = rs.size - 1
    bs = np.zeros(nrows + 2, dtype=np.int32)
    nb = C.c_int32()
    checkh(libspxhost().spxh_block_rows(nrows, rs.ctypes.data, cap, C.byref(nb), bs.ctypes.data),
           "spxh_block_rows")
    return bs[: nb.value + 1]


class Ogs:
    """ogs_t set up on one rank from an id list (ogsSetup.cpp:62-71)."""

    def __init__(self, ids, kind=_lib.SIGNED, unique=True, verbose=False, device_id=0):
        platform_open(device_id)
        ids = np.ascontiguousarray(ids, dtype=np.int64).copy()
        self.h = C.c_void_p()
        checkh(libspxhost().spxh_ogs_setup(ids.size, ids.ctypes.data, kind, 1 if unique else 0,
                                           1 if verbose else 0, C.byref(self.h)), "spxh_ogs_setup")
        self.ids = ids  # sign-flagged when unique
        self.refresh()

    @classmethod
    def hex(cls, nx, ny, nz, p, kind=_lib.SIGNED, unique=True, keep_ids=True, device_id=0):
        """ogs_t of the hex box's lexicographic GLL ids, built in closed form (the same operator and
        flags as Ogs(hex_ids(nx, ny, nz, p)))."""
        platform_open(device_id)
        self = cls.__new__(cls)
        ids = np.empty(nx * ny * nz * (p + 1) ** 3, dtype=np.int64) if keep_ids else None
        self.h = C.c_void_p()
        checkh(libspxhost().spxh_ogs_setup_hex(nx, ny, nz, p, kind, 1 if unique else 0,
                                               ids.ctypes.data if keep_ids else None, C.byref(self.h)),
               "spxh_ogs_setup_hex")
        self.ids = ids
        self.refresh()
        return self

    @classmethod
    def box(cls, element, nx, ny, nz, p, kind=_lib.SIGNED, unique=True, keep_ids=True, device_id=0):
        """mesh_t's ogs_t of a box of `element` ("Tri", "Quad", "Tet", "Hex"): the operator of
        Ogs(box_ids(...)) with the gather+scatter tile order the mesh uses."""
        platform_open(device_id)
        self = cls.__new__(cls)
        ne, npe = box_shape(element, nx, ny, nz, p)
        ids = np.empty(ne * npe, dtype=np.int64) if keep_ids else None
        self.h = C.c_void_p()
        checkh(libspxhost().spxh_ogs_setup_box(ELEMENT_TYPES[element], nx, ny, nz, p, kind, 1 if unique else 0,
                                               ids.ctypes.data if keep_ids else None, C.byref(self.h)),
               "spxh_ogs_setup_box")
        self.ids = ids
        self.refresh()
        return self

    @classmethod
    def batch(cls, parts):
        """The block-diagonal union of set-up handles: part i's local nodes follow parts 0..i-1's."""
        self = cls.__new__(cls)
        arr = (C.c_void_p * len(parts))(*[p.h.value for p in parts])
        self.h = C.c_void_p()
        checkh(libspxhost().spxh_ogs_setup_batch(arr, len(parts), C.byref(self.h)), "spxh_ogs_setup_batch")
        self.ids = None
        self.refresh()
        return self

    def refresh(self):
        info = _lib.OgsInfo()
        checkh(libspxhost().spxh_ogs_info(self.h, C.byref(info)), "spxh_ogs_info")
        self.info = {f: getattr(info, f) for f, _ in _lib.OgsInfo._fields_}
        self.info["mixed_tiles"] = list(self.info["mixed_tiles"])
        for k in ("N", "Ngather", "NlocalT", "NlocalP", "NgatherGlobal", "gather_defined"):
            setattr(self, k, self.info[k])

    def csr(self, which="T"):
        w = 0 if which == "T" else 1
        rows = self.info["NlocalT"] if w == 0 else self.info["NlocalP"]
        nnz = self.info["nnzT"] if w == 0 else self.info["nnzN"]
        nb = self.info["NblocksT"] if w == 0 else self.info["NblocksN"]
        rs = np.zeros(rows + 1, dtype=np.int32)
        ci = np.zeros(max(nnz, 1), dtype=np.int32)
        bs = np.zeros(nb + 1, dtype=np.int32)
        checkh(libspxhost().spxh_ogs_copy_csr(self.h, w, rs.ctypes.data, ci.ctypes.data, bs.ctypes.data),
               "spxh_ogs_copy_csr")
        return rs, ci[:nnz], bs

    def set_block_nnz(self, cap, gs_cap=None):
        """Row-block LDS tiles: gather/scatter blocks get `cap`, gather+scatter blocks `gs_cap`
        (default: the same)."""
        checkh(libspxhost().spxh_ogs_set_block_nnz(self.h, cap), "spxh_ogs_set_block_nnz")
        checkh(libspxhost().spxh_ogs_set_gs_block_nnz(self.h, gs_cap or cap), "spxh_ogs_set_gs_block_nnz")
        self.refresh()

    def set_gs_format(self, sorted_runs, xcd_remap=-1):
        """Sym gather+scatter on the sorted-run row blocks (True, default) or the T-CSR blocks (False);
        xcd_remap: sorted-run launch hint (-1 library default)."""
        checkh(libspxhost().spxh_ogs_set_gs_format(self.h, 1 if sorted_runs else 0, xcd_remap),
               "spxh_ogs_set_gs_format")

    def set_gs_nontemporal(self, mode=-1):
        """Non-temporal value loads and stores of the Sym gather+scatter: -1 where the tiles' builder
        set them (node sizes whose 128-byte lines lie whole in one tile), 0 never, 1 always.  Results
        do not change."""
        checkh(libspxhost().spxh_ogs_set_gs_nontemporal(self.h, mode), "spxh_ogs_set_gs_nontemporal")
        self.refresh()

    def set_gs_row_keys(self, keys=None):
        """Rows of the sorted-run gather+scatter tiles sorted by keys (one int64 per T-CSR row, ties
        in row order); None restores the CSR's order.  Results do not change."""
        if keys is None:
            checkh(libspxhost().spxh_ogs_set_gs_row_keys(self.h, None), "spxh_ogs_set_gs_row_keys")
        else:
            k = np.ascontiguousarray(keys, dtype=np.int64)
            if k.size != self.NlocalT:
                raise ValueError(f"need {self.NlocalT} keys, got {k.size}")
            checkh(libspxhost().spxh_ogs_set_gs_row_keys(self.h, k.ctypes.data), "spxh_ogs_set_gs_row_keys")
        self.refresh()

    def set_gs_tiles(self, keys):
        """Rows of the sorted-run gather+scatter tiles sorted by keys, and no tile holding rows of two
        keys.  Results do not change."""
        k = np.ascontiguousarray(keys, dtype=np.int64)
        if k.size != self.NlocalT:
            raise ValueError(f"need {self.NlocalT} keys, got {k.size}")
        checkh(libspxhost().spxh_ogs_set_gs_tiles(self.h, k.ctypes.data), "spxh_ogs_set_gs_tiles")
        self.refresh()

    def set_gs_mixed_tiles(self, storage, keys=None):
        """As set_gs_tiles for the FP32 (storage STORE_F32) or BF16 (STORE_BF16) gather+scatter alone;
        None returns it to the tiles of the other types.  Results do not change."""
        if keys is None:
            checkh(libspxhost().spxh_ogs_set_gs_mixed_tiles(self.h, storage, None), "spxh_ogs_set_gs_mixed_tiles")
        else:
            k = np.ascontiguousarray(keys, dtype=np.int64)
            if k.size != self.NlocalT:
                raise ValueError(f"need {self.NlocalT} keys, got {k.size}")
            checkh(libspxhost().spxh_ogs_set_gs_mixed_tiles(self.h, storage, k.ctypes.data),
                   "spxh_ogs_set_gs_mixed_tiles")
        self.refresh()

    def set_gather_format(self, gather_sorted=None, scatter_sorted=None):
        """T-CSR gather (Trans) and scatter (NoTrans) on the sorted-run blocks (True) or the T-CSR row
        blocks (False); None keeps the current choice (both sorted by default)."""
        enc = lambda x: -1 if x is None else (1 if x else 0)
        checkh(libspxhost().spxh_ogs_set_gather_format(self.h, enc(gather_sorted), enc(scatter_sorted)),
               "spxh_ogs_set_gather_format")

    def autotune(self, dtype=np.float64, k=1, reps=5, mixed=False):
        """Time every launch shape of the Sym gather+scatter for k fields of dtype (mixed=True: the
        FP64-sum kernel on float32 / uint16-bf16 storage), keep the fastest, return it."""
        best = _lib.GsTuning()
        if mixed:
            rc = libspxhost().spxh_ogs_autotune(self.h, 0, _lib.storage_of(dtype), k, reps, C.byref(best))
        else:
            rc = libspxhost().spxh_ogs_autotune(self.h, ogs_type(np.dtype(dtype)), -1, k, reps, C.byref(best))
        checkh(rc, "spxh_ogs_autotune")
        self.refresh()
        return {f: getattr(best, f) for f, _ in _lib.GsTuning._fields_}

    def set_scatter_path(self, by_rows):
        """Scatter over the CSR row blocks (True, default) or through the local->row map (False)."""
        checkh(libspxhost().spxh_ogs_set_scatter_path(self.h, 1 if by_rows else 0), "spxh_ogs_set_scatter_path")

    def global_to_local(self):
        """SetupGlobalToLocalMapping (ogsSetup.cpp:833-849): the T-row of every local node, -1 where
        the id is zero."""
        g = np.empty(self.N, dtype=np.int32)
        checkh(libspxhost().spxh_ogs_global_to_local(self.h, g.ctypes.data), "spxh_ogs_global_to_local")
        return g

    def gather_scatter(self, q, k=1, op=_lib.ADD, trans=_lib.SYM):
        checkh(libspxhost().spxh_ogs_gather_scatter(self.h, ogs_type(q.dtype), op, trans, k, q.ptr),
               "spxh_ogs_gather_scatter")

    def gather_scatter_mixed(self, q, k=1):
        """Sym Add gatherScatter of an FP32 or BF16 (uint16 bits) vector, FP64 row sums."""
        checkh(libspxhost().spxh_ogs_gather_scatter_mixed(self.h, _lib.storage_of(q.dtype), k, q.ptr),
               "spxh_ogs_gather_scatter_mixed")

    def gather(self, gq, q, k=1, op=_lib.ADD, trans=_lib.TRANS):
        checkh(libspxhost().spxh_ogs_gather(self.h, ogs_type(q.dtype), op, trans, k, q.ptr, gq.ptr),
               "spxh_ogs_gather")

    def scatter(self, q, gq, k=1, trans=_lib.NOTRANS):
        checkh(libspxhost().spxh_ogs_scatter(self.h, ogs_type(q.dtype), trans, k, gq.ptr, q.ptr),
               "spxh_ogs_scatter")

    def __del__(self):
        try:
            if self.h.value:
                libspxhost().spxh_ogs_destroy(self.h)
        except Exception:
            pass
