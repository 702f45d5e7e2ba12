"""The product's host setup (libspxhost, no GPU) against the oracle's restatement."""
import numpy as np
import pytest

import streamparanumal_amd as spx
from streamparanumal_amd import ogs
from streamparanumal_amd._lib import SpxError
from oracle import pyoracle as o


@pytest.mark.parametrize("mesh", [(1, 1, 1, 1), (2, 2, 1, 2), (3, 2, 2, 3), (5, 3, 2, 4), (8, 8, 8, 7), (4, 4, 4, 15)])
def test_hex_ids_and_csr_match_oracle(mesh):
    ids = ogs.hex_ids(*mesh)
    assert np.array_equal(ids, o.hex_ids(*mesh))
    flagged, csr = ogs.build_csr(ids, spx.SIGNED, True)
    ref = o.Ogs(ids, kind=1, unique=True)
    assert csr["NrowsT"] == ref.NrowsT and csr["NrowsN"] == ref.NrowsN
    for k in ("rowStartsT", "colIdsT", "rowStartsN", "colIdsN"):
        assert np.array_equal(csr[k], getattr(ref, k)), k
    assert np.array_equal(flagged, ref.ids)
    assert csr["gather_defined"]


def test_min_label_ids_give_same_operator():
    mesh = (3, 3, 2, 4)
    _, a = ogs.build_csr(o.hex_ids(*mesh, scheme="lex"))
    _, b = ogs.build_csr(o.hex_ids(*mesh, scheme="min"))
    for k in ("rowStartsT", "colIdsT"):
        assert np.array_equal(a[k], b[k])


@pytest.mark.parametrize("kind,unique", [(spx.SIGNED, False), (spx.UNSIGNED, False), (spx.SIGNED, True)])
def test_random_id_lists(kind, unique):
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(1, 600))
        hi = int(rng.choice([5, 50, 10 ** 12]))  # dense and sparse label ranges
        ids = rng.integers(-hi, hi + 1, size=n).astype(np.int64)
        if unique:
            ids = np.abs(ids)
        flagged, csr = ogs.build_csr(ids, kind, unique)
        ref = o.Ogs(ids, kind=kind, unique=unique)
        assert ref.error == 0
        assert csr["NrowsT"] == ref.NrowsT and csr["NrowsN"] == ref.NrowsN
        for k in ("rowStartsT", "colIdsT", "rowStartsN", "colIdsN"):
            assert np.array_equal(csr[k], getattr(ref, k)), (trial, k)
        assert csr["gather_defined"] == ref.gather_defined
        assert np.array_equal(flagged, ref.ids)


def test_setup_rejects_invalid_requests():
    ids = np.arange(1, 11, dtype=np.int64)
    with pytest.raises(spx.SpxError):
        ogs.build_csr(ids, spx.UNSIGNED, True)
    with pytest.raises(spx.SpxError):
        ogs.build_csr(ids, spx.HALO, False)


def test_empty_and_all_zero_ids():
    for ids in (np.zeros(0, np.int64), np.zeros(7, np.int64)):
        _, csr = ogs.build_csr(ids, spx.SIGNED, True)
        assert csr["NrowsT"] == 0 and csr["rowStartsT"].tolist() == [0]


@pytest.mark.parametrize("cap", [8, 512, 2048, 4096])
def test_greedy_row_blocks(cap):
    _, csr = ogs.build_csr(ogs.hex_ids(6, 5, 4, 7))
    rs = csr["rowStartsT"]
    bs = ogs.block_rows(rs, cap)
    assert bs[0] == 0 and bs[-1] == rs.size - 1 and np.all(np.diff(bs) > 0)
    nnz = rs[bs[1:]] - rs[bs[:-1]]
    assert np.all(nnz <= cap)
    # greedy: a block cannot take the next row without overflowing
    nxt = rs[np.minimum(bs[1:-1] + 1, rs.size - 1)] - rs[bs[:-2]]
    assert np.all(nxt > cap)


def test_row_larger_than_cap_is_refused():
    _, csr = ogs.build_csr(ogs.hex_ids(2, 2, 2, 1))
    with pytest.raises(spx.SpxError):
        ogs.block_rows(csr["rowStartsT"], 4)


# ---------------------------------------------------------------------------
# sorted-run row blocks: decoded on the CPU and compared with the CSR
# ---------------------------------------------------------------------------

def decode_srb(s):
    """(local index, previous-member position, row if row-leading) of every entry of every tile, by
    the device kernels' decoding rule, tile by tile in position order."""
    cpt = s["chunks_per_tile"]
    tiles = []
    for b in range(s["nblocks"]):
        ent = []
        for k in range(cpt):
            C = s["chunks"][b * cpt + k]
            rm, lm = int(C["runMask"]), int(C["linkMask"])
            for l in range(int(C["nvalid"])):
                run = bin(rm & ((2 << l) - 1)).count("1") - 1
                idx = int(s["runIds"][C["runBase"] + run]) + l
                before = bin(lm & ((1 << l) - 1)).count("1")
                prev = int(s["links"][C["linkBase"] + before]) if lm >> l & 1 else -1
                row = int(C["chunkRow"]) + (l - before) if prev < 0 else None
                ent.append((idx, prev, row))
            # no bits past the entries (a padding record keeps run bit 0, so it still decodes)
            assert all(not (rm >> l & 1) and not (lm >> l & 1) for l in range(max(int(C["nvalid"]), 1), 64))
        tiles.append(ent)
    return tiles


@pytest.mark.parametrize("mesh,cap", [((2, 2, 2, 2), 64), ((3, 2, 1, 3), 100), ((4, 3, 2, 4), 4096),
                                      ((5, 5, 5, 1), 256), ((2, 2, 2, 7), 1000)])
def test_srb_layout_matches_csr(mesh, cap):
    ids = ogs.hex_ids(*mesh)
    _, c = ogs.build_csr(ids)
    rs, ci = c["rowStartsT"], c["colIdsT"]
    bs = ogs.block_rows(rs, cap)
    s = ogs.srb_build(rs, ci, bs, cap)
    cpt = s["chunks_per_tile"]
    assert cpt >= 4 and cpt & (cpt - 1) == 0 and 64 * cpt >= cap          # power of two, covers the cap
    assert s["nblocks"] == bs.size - 1 and s["nnz"] == ci.size and s["nchunks"] == s["nblocks"] * cpt
    assert np.array_equal(s["entryStarts"], rs[bs])
    assert s["nlinks"] == ci.size - (rs.size - 1)            # every member but the first of each row
    assert s["rows_ascending"]                                # rows numbered by their leading positions
    for b, ent in enumerate(decode_srb(s)):
        e0, e1 = rs[bs[b]], rs[bs[b + 1]]
        n = e1 - e0
        assert len(ent) == n
        dec = [e[0] for e in ent]
        assert dec == sorted(int(x) for x in ci[e0:e1])        # ascending local index
        # the links from each row's first member walk the CSR row in order; rows in CSR order
        nxt = {}
        for i, (_, prev, _) in enumerate(ent):
            if prev >= 0:
                assert prev < i and prev not in nxt
                nxt[prev] = i
        lead = [i for i, e in enumerate(ent) if e[1] < 0]
        assert [ent[i][2] for i in lead] == list(range(bs[b], bs[b + 1]))
        for r, i in zip(range(bs[b], bs[b + 1]), lead):
            walk = [ent[i][0]]
            while i in nxt:
                i = nxt[i]
                walk.append(ent[i][0])
            assert walk == [int(x) for x in ci[rs[r]:rs[r + 1]]]
    for rec in s["chunks"][: s["nchunks"]]:
        assert int(rec["runMask"]) & 1                       # every record decodes (padding included)
        assert 0 <= rec["runBase"] < max(s["nruns"], 1)


def signed_ids(n, seed):
    """Random ids with repeats, negatives (groups with no positive member among them) and zeros."""
    ids = (o.rand_int(n, 300 + seed) % (n // 3) + 1).astype(np.int64)
    r = o.rand_int(n, 400 + seed)
    ids[r % 4 == 0] *= -1
    ids[r % 11 == 0] = 0
    return ids


@pytest.mark.parametrize("seed,unique,cap", [(0, False, 64), (1, False, 700), (2, True, 256), (3, True, 4096)])
def test_srb_member_masks(seed, unique, cap):
    """Bit l of record k's mask is set exactly when position l holds a member of the N-CSR."""
    n = 3000
    _, c = ogs.build_csr(signed_ids(n, seed), spx.SIGNED, unique)
    rs, ci = c["rowStartsT"], c["colIdsT"]
    if unique:                                             # one positive member per group
        assert c["NrowsN"] == c["NrowsT"] and np.all(np.diff(c["rowStartsN"]) == 1)
    else:                                                  # groups with none, and with several
        assert c["NrowsN"] < c["NrowsT"] and np.diff(c["rowStartsN"]).max() > 1
    bs = ogs.block_rows(rs, cap)
    s = ogs.srb_build(rs, ci, bs, cap, n_csr=(c["rowStartsN"], c["colIdsN"]), ncols=n)
    assert s["nrowsN"] == c["NrowsN"]
    # without unique flagging the groups with no positive member follow the others (ogsSetup.cpp:517-546),
    # so the tile where they start holds rows out of first-member order: no row numbers on this layout
    first = ci[rs[:-1]]
    assert s["rows_ascending"] == bool(np.all(np.diff(first) > 0)) == unique
    members = set(int(x) for x in c["colIdsN"])
    cpt = s["chunks_per_tile"]
    for b, ent in enumerate(decode_srb(s)):
        for k in range(cpt):
            m = int(s["memberN"][b * cpt + k])
            chunk = ent[64 * k: 64 * (k + 1)]
            assert m >> len(chunk) == 0                    # no bits past the entries
            assert [bool(m >> l & 1) for l in range(len(chunk))] == [e[0] in members for e in chunk]
    plain = ogs.srb_build(rs, ci, bs, cap)
    assert np.array_equal(plain["chunks"], s["chunks"]) and np.array_equal(plain["links"], s["links"])


def test_srb_signed_rejects_misaligned_n_csr():
    """N-CSR rows that are not the members of the same T-CSR rows are refused (the kernels would
    fold the wrong rows)."""
    rs = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 2, 3], np.int32)
    bs = np.array([0, 2], np.int32)
    ok = ogs.srb_build(rs, ci, bs, 64, n_csr=(np.array([0, 1, 3], np.int32), np.array([1, 2, 3], np.int32)),
                       ncols=4)
    assert int(ok["memberN"][0]) == 0b1110
    bad = [((0, 1, 2), (2, 0)),        # rows swapped
           ((0, 1), (2,)),             # row 0's N member lies in T row 1
           ((0, 2), (1, 0)),           # members out of order
           ((0, 0, 1), (3,)),          # an empty N row, then a member of T row 1 in N row 1 -- allowed
           ((0, 1, 1, 1), (0,))]       # more N rows than T rows
    for i, (rsn, cin) in enumerate(bad):
        args = (np.array(rsn, np.int32), np.array(cin, np.int32))
        if i == 3:
            assert int(ogs.srb_build(rs, ci, bs, 64, n_csr=args, ncols=4)["memberN"][0]) == 0b1000
            continue
        with pytest.raises(SpxError):
            ogs.srb_build(rs, ci, bs, 64, n_csr=args, ncols=4)


def test_srb_rows_ascending_flag():
    """The layout numbers rows by their leading positions, so it reports whether each tile's rows
    ascend by first member (and have one): the gather and scatter need it, the gather+scatter not."""
    bs = np.array([0, 2], np.int32)
    up = ogs.srb_build(np.array([0, 2, 3], np.int32), np.array([0, 2, 1], np.int32), bs, 64)
    assert up["rows_ascending"]                                   # first members 0 < 1
    down = ogs.srb_build(np.array([0, 1, 3], np.int32), np.array([2, 0, 1], np.int32), bs, 64)
    assert not down["rows_ascending"]                             # first members 2 > 0
    split = ogs.srb_build(np.array([0, 1, 3], np.int32), np.array([2, 0, 1], np.int32),
                          np.array([0, 1, 2], np.int32), 64)
    assert split["rows_ascending"]                                # one row per tile: nothing to order
    empty = ogs.srb_build(np.array([0, 2, 2], np.int32), np.array([0, 1], np.int32), bs, 64)
    assert not empty["rows_ascending"]                            # an empty row has no leading position


def test_srb_rejects_bad_blocks():
    rs = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 1, 2], np.int32)
    with pytest.raises(SpxError):
        ogs.srb_build(rs, ci, np.array([0, 2], np.int32), 2)            # a block of 3 entries, cap 2
    with pytest.raises(SpxError):
        ogs.srb_build(rs, ci, np.array([0, 1], np.int32), 4)            # blocks stop short of the rows
    with pytest.raises(SpxError):
        ogs.srb_build(rs, np.array([0, 0, 2], np.int32), np.array([0, 2], np.int32), 4)  # node twice
    empty = ogs.srb_build(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(1, np.int32), 16)
    assert empty["nblocks"] == 0 and empty["nnz"] == 0


# ---------------------------------------------------------------------------
# batched instances: the generator is the reference LCG stream
# ---------------------------------------------------------------------------

def test_batch_instances_follow_the_lcg():
    from streamparanumal_amd.batch import instances, lcg_uniform01
    u = np.array(lcg_uniform01(200, 4))
    assert np.array_equal(u, (o.uniform(200, 4) + 1.0) / 2.0)     # U[-1,1) draw -> U[0,1), exactly
    inst = instances(64, 4)
    assert len(inst) == 64
    assert all(4 <= n <= 48 and p in (3, 5, 7, 9) for n, p in inst)
    assert inst[:4] == [(4 + int(u[0] * 45), (3, 5, 7, 9)[int(u[1] * 4)]),
                        (4 + int(u[2] * 45), (3, 5, 7, 9)[int(u[3] * 4)]),
                        (4 + int(u[4] * 45), (3, 5, 7, 9)[int(u[5] * 4)]),
                        (4 + int(u[6] * 45), (3, 5, 7, 9)[int(u[7] * 4)])]
    # every degree and both ends of the range occur over the 64 draws
    assert {p for _, p in inst} == {3, 5, 7, 9}


# ---------------------------------------------------------------------------
# closed-form hex-box operator: identical to the generic setup over the ids
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("box", [(1, 1, 1, 1), (2, 1, 1, 1), (2, 3, 4, 2), (3, 2, 2, 5), (1, 4, 2, 7),
                                 (5, 5, 5, 1), (2, 2, 2, 15), (6, 5, 4, 3)])
@pytest.mark.parametrize("kind,unique", [(spx.SIGNED, True), (spx.SIGNED, False), (spx.UNSIGNED, False)])
def test_hex_box_setup_equals_generic(box, kind, unique):
    ids0, c0 = ogs.build_csr(ogs.hex_ids(*box), kind, unique)
    ids1, c1 = ogs.build_csr_hex(*box, kind, unique)
    assert np.array_equal(ids0, ids1)                      # the same flags on the same nodes
    assert c0.keys() == c1.keys()
    for k in c0:
        if isinstance(c0[k], np.ndarray):
            assert np.array_equal(c0[k], c1[k]), k
        else:
            assert c0[k] == c1[k], k


def test_hex_box_setup_rejects():
    with pytest.raises(SpxError):
        ogs.build_csr_hex(0, 1, 1, 3)
    with pytest.raises(SpxError):
        ogs.build_csr_hex(1, 1, 1, 0)
    with pytest.raises(SpxError):
        ogs.build_csr_hex(2, 2, 2, 3, spx.UNSIGNED, True)    # Unsigned cannot be unique (ogsSetup.cpp:92)


def test_hex_box_setup_with_a_smaller_thread_team():
    """OMP_THREAD_LIMIT below the requested team: the closed-form setup's scans split the rows by
    the team the runtime gives, so the operator is still the generic setup's."""
    import os
    import subprocess
    import sys
    code = ("import numpy as np, streamparanumal_amd as spx; from streamparanumal_amd import ogs\n"
            "box = (6, 5, 4, 3)\n"
            "i0, c0 = ogs.build_csr(ogs.hex_ids(*box), spx.SIGNED, True)\n"
            "i1, c1 = ogs.build_csr_hex(*box, spx.SIGNED, True)\n"
            "assert np.array_equal(i0, i1)\n"
            "assert all(np.array_equal(c0[k], c1[k]) if isinstance(c0[k], np.ndarray) else c0[k] == c1[k] for k in c0)\n"
            "print('ok')\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, OMP_NUM_THREADS="8", OMP_THREAD_LIMIT="3", PYTHONPATH=root)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]


@pytest.mark.parametrize("box,r,cap", [((6, 5, 4, 7), 1, 2048), ((6, 5, 4, 7), 2, 2048), ((6, 5, 4, 7), 4, 4096),
                                       ((9, 3, 2, 3), 2, 2048), ((4, 4, 4, 8), 2, 2048), ((20, 3, 3, 7), 2, 2048),
                                       ((64, 3, 2, 7), 2, 2048), ((6, 3, 2, 15), 1, 2048),
                                       ((30, 2, 2, 15), 1, 2048)])
def test_hex_line_tiles(box, r, cap):
    """Line-aligned gather+scatter tiles: a permutation of the rows cut into tiles of at most cap
    members, in which every group of r node rows of an element (one memory line of the vector) lies
    in one tile, except where a group of y points too large for one tile is cut along x."""
    nx, ny, nz, p = box
    ids, c = ogs.build_csr_hex(nx, ny, nz, p, spx.SIGNED, True)
    rs, ci = c["rowStartsT"], c["colIdsT"]
    nrows = rs.size - 1
    grid = np.abs(ids[ci[rs[:-1]]]) - 1
    order, cuts = ogs.hex_gs_tiles(nx, ny, nz, p, r, cap, grid)
    assert np.array_equal(np.sort(order), np.arange(nrows)) and cuts[0] == 1
    tile_of_pos = np.cumsum(cuts) - 1                      # tile of each entry of order
    tile_of_row = np.empty(nrows, dtype=np.int64)
    tile_of_row[order] = tile_of_pos
    sizes = np.bincount(tile_of_pos, weights=np.diff(rs)[order])
    row_of = np.empty(ids.size, dtype=np.int64)
    row_of[ci] = np.repeat(np.arange(nrows), np.diff(rs))
    Nq = p + 1
    t = tile_of_row[row_of].reshape(nx * ny * nz, Nq, Nq, Nq)         # element, k, j, i
    split = 0
    for j0 in range(0, Nq, r):
        g = t[:, :, j0:j0 + r, :].reshape(nx * ny * nz, Nq, -1)
        split += int((g != g[..., :1]).any(axis=-1).sum())
    G = nx * p + 1
    big = (nx * (p + 1)) * 2 * r * 2 > cap                   # some y group may exceed cap: x cuts allowed
    assert sizes.max() <= cap or big
    if not big:
        assert split == 0
    else:
        assert 0 < split <= (nx * ny * nz * Nq * Nq // r) // 16   # only the element columns at x cuts


def whole_line_mask(ci, entry_starts):
    """spx_srb_t.nontemporal from its definition: bit e when at least 85 % of the (tile, 128-byte line)
    pairs of a vector of 2^e-byte nodes that hold an entry hold all 2^(7-e) nodes of the line."""
    tile = np.repeat(np.arange(entry_starts.size - 1), np.diff(entry_starts))
    ids = ci[: entry_starts[-1]].astype(np.int64)
    mask = 0
    for e in range(1, 6):
        sh = 7 - e
        line = ids >> sh
        _, counts = np.unique(tile * (int(line.max()) + 1) + line, return_counts=True)
        if np.count_nonzero(counts == 1 << sh) >= 0.85 * counts.size:
            mask |= 1 << e
    return mask


def tiles_from_order(rs, ci, order, cuts, cap):
    """The CSR with its rows in `order`, and greedy row blocks of at most cap entries that also start at
    every cut (blockRowsGreedy with cuts)."""
    sizes = np.diff(rs)[order]
    rs2 = np.zeros(order.size + 1, dtype=np.int64)
    rs2[1:] = np.cumsum(sizes)
    ci2 = np.concatenate([ci[rs[r]:rs[r + 1]] for r in order])
    bs, acc = [0], 0
    for i, sz in enumerate(sizes):
        if i > 0 and (cuts[i] or acc + sz > cap):
            bs.append(i)
            acc = 0
        acc += sz
    bs.append(order.size)
    return rs2, ci2, np.asarray(bs)


@pytest.mark.parametrize("case", ["line2", "line1", "blocks_p1", "random"])
def test_srb_nontemporal_mask(case):
    """The layout builder's non-temporal mask equals its definition on hex line tiles of 2 node rows at
    p = 7 (whole 128-byte lines of FP64 nodes: bit 3), 1-row line tiles, lattice blocks at p = 1 and a
    CSR of random ids (short runs: no bit)."""
    cap = 2048
    if case == "random":
        ids = (o.rand_int(30000, 7) % 9000 + 1).astype(np.int64)
        _, c = ogs.build_csr(ids)
        rs, ci = c["rowStartsT"], c["colIdsT"]
        bs = ogs.block_rows(rs, cap)
    else:
        nx, ny, nz, p = (64, 3, 2, 7) if case.startswith("line") else (24, 8, 8, 1)
        ids, c = ogs.build_csr_hex(nx, ny, nz, p, spx.SIGNED, True)
        rs, ci = c["rowStartsT"], c["colIdsT"]
        grid = np.abs(ids[ci[rs[:-1]]]) - 1
        if case.startswith("line"):
            order, cuts = ogs.hex_gs_tiles(nx, ny, nz, p, 2 if case == "line2" else 1, cap, grid)
        else:
            Mx, My = nx * p + 1, ny * p + 1                  # 8 x 4 x 4 blocks of lattice points
            i, j, k = grid % Mx, (grid // Mx) % My, grid // (Mx * My)
            key = ((k // 4) * (-(-My // 4)) + j // 4) * (-(-Mx // 8)) + i // 8
            order = np.argsort(key, kind="stable").astype(np.int32)
            cuts = np.zeros(order.size, np.uint8)
        rs, ci, bs = tiles_from_order(rs, ci, order, cuts, cap)
    srb = ogs.srb_build(rs, ci, bs, cap)
    want = whole_line_mask(ci, srb["entryStarts"])
    assert srb["nontemporal"] == want
    if case == "line2":
        assert want & 0x08                        # FP64 nodes: whole lines
    if case in ("blocks_p1", "random"):
        assert want & 0x0E == 0                   # 2-, 4- and 8-byte nodes: split lines
