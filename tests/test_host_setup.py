"""The product's host setup (libspxhost, no GPU) against the oracle's restatement."""
import numpy as np
import pytest

import streamparanumal_amd as spx
from streamparanumal_amd import ogs
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
