"""Pins of the CPU oracle (CPU only).

The reference holds no golden vectors, known-answer tests or fixtures for
this path (it never initialises or checks data).  Each oracle function is
pinned here by something other than its own formula: big-int recurrences,
exact integer arithmetic, hand-derived order probes whose expected values
follow from IEEE round-to-nearest-even and fail for any other summation
order, brute force on tiny meshes, and closed-form counts.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import pyoracle as o

A, Cc, M = 6364136223846793005, 1, 1 << 64


def lcg_python(seed, n):
    x = (A * ((M - 1) ^ seed) + Cc) % M
    out = []
    for _ in range(n):
        out.append(x)
        x = (A * x + Cc) % M
    return out


# ------------------------------------------------------------------ LCG ----
@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 12345])
def test_lcg_states_match_bigint_recurrence(seed):
    xs = lcg_python(seed, 1100)
    for k in (0, 1, 2, 7, 64, 511, 512, 1099):
        assert o.lcg_state(seed, k) == xs[k]


def test_uniform_mapping_exact():
    xs = lcg_python(1, 2000)
    v = o.uniform(2000, 1)
    for x, got in zip(xs, v):
        want = Fraction(-1) + Fraction(x >> 11, 1 << 52)
        assert Fraction(got) == want
    assert v.min() >= -1.0 and v.max() < 1.0
    # continuing a stream = skipping ahead
    assert np.array_equal(o.uniform(500, 1, 1500), v[1500:])


def test_rand_int_is_abs_of_low_word():
    xs = lcg_python(0, 1000)
    got = o.rand_int(1000, 0)
    for x, g in zip(xs, got):
        w = x & 0xFFFFFFFF
        s = w - (1 << 32) if w & 0x80000000 else w
        assert g == abs(s)


def test_loguniform_draw_order():
    """Element i: magnitude 10^(-6u) from draw 2(skip+i), u = (x >> 11) 2^-53, sign from the top bit
    of draw 2(skip+i)+1 -- recomputed from the Python LCG with the same libm pow."""
    import math
    n, seed, skip = 500, 3, 17
    xs = lcg_python(seed, 2 * (skip + n))
    want = np.empty(n)
    for i in range(n):
        u = (xs[2 * (skip + i)] >> 11) * 2.0 ** -53
        mag = math.pow(10.0, -6.0 * u)
        want[i] = -mag if xs[2 * (skip + i) + 1] >> 63 else mag
    assert np.array_equal(o.loguniform(n, seed, skip).view(np.uint64), want.view(np.uint64))


def test_loguniform_range_and_signs():
    v = o.loguniform(20000, 3)
    a = np.abs(v)
    assert a.min() >= 1e-6 * (1 - 1e-12) and a.max() <= 1.0
    assert 0.45 < np.mean(v < 0) < 0.55
    # magnitude decades are uniform: each of the 6 decades holds ~1/6
    dec = np.floor(-np.log10(a)).clip(0, 5)
    counts = np.bincount(dec.astype(int), minlength=6) / a.size
    assert np.all(np.abs(counts - 1 / 6) < 0.02)


# ------------------------------------------------------------- BS1, BS2 ----
def test_copy_is_identity():
    a = o.uniform(12345, 1)
    assert np.array_equal(o.bs1_copy(a).view(np.uint64), a.view(np.uint64))


def test_axpy_exact_on_dyadic_inputs():
    k = np.arange(-500, 500)
    x, y = k / 8.0, (k[::-1] % 17) / 4.0
    got = o.bs2_axpy(0.5, x, -2.0, y)
    want = [float(Fraction(1, 2) * Fraction(a) - 2 * Fraction(b)) for a, b in zip(x, y)]
    assert np.array_equal(got, np.array(want))


def test_axpy_is_not_contracted():
    # alpha*x = 1 + 2^-29 + 2^-60 rounds to 1 + 2^-29; beta*y = -(1 + 2^-29) exactly.
    # Two roundings and one add give 0; a fused multiply-add would give 2^-60.
    a = 1.0 + 2.0 ** -30
    got = o.bs2_axpy(a, np.array([a]), -1.0, np.array([1.0 + 2.0 ** -29]))
    assert got[0] == 0.0


# -------------------------------------------------------- BS3..BS5 order ----
def test_norm_integer_known_answer():
    rng = np.random.default_rng(7)
    a = rng.integers(-1000, 1000, size=300001).astype(np.float64)
    assert o.bs3_norm(a) == float(sum(int(x) ** 2 for x in a))


def test_dot_serial_chain_probe():
    # N = 196608, bs = 256: Nblock = 256, stride 65536.  Thread (0,0) sums a[0], a[65536],
    # a[131072] in that order: 2^53 + 1 -> 2^53 (tie to even), + 1 -> 2^53.  Any order that adds
    # the two ones first gives 2^53 + 2.
    n = 196608
    a, b = np.zeros(n), np.ones(n)
    a[0], a[65536], a[131072] = 2.0 ** 53, 1.0, 1.0
    assert o.bs4_dot(a, b) == 2.0 ** 53


def test_dot_block_tree_probe():
    # N = 256: one block.  Tree step w=128 adds s[128] into s[0]: 1 + 2^53 -> 2^53; s[1] = 1
    # survives to the last step s[0]+s[1] = 2^53 + 1 -> 2^53.  Left-to-right or pairwise sums
    # give 1 + 1 + 2^53 = 2^53 + 2.
    a, b = np.zeros(256), np.ones(256)
    a[0], a[1], a[128] = 1.0, 1.0, 2.0 ** 53
    assert o.bs4_dot(a, b) == 2.0 ** 53
    assert sum([1.0, 1.0, 2.0 ** 53]) == 2.0 ** 53 + 2


def test_second_stage_probe():
    # N = 512, bs = 256: Nblock = 2 and thread t of block b holds a[t + 256 b].  Block 0's tree
    # gives partial[0] = 2^53 (from a[128]); block 1 gives partial[1] = a[256] + a[257].
    # Stage 2 adds them as s[0] + s[1]: 2^53 + 1 -> 2^53, but 2^53 + 2 is exact.
    a, b = np.zeros(512), np.ones(512)
    a[128], a[256] = 2.0 ** 53, 1.0
    assert o.bs4_dot(a, b) == 2.0 ** 53
    a[257] = 1.0
    assert o.bs4_dot(a, b) == 2.0 ** 53 + 2


@pytest.mark.parametrize("n", [1, 255, 256, 257, 65536, 65537, 1000003])
def test_reductions_close_to_exact(n):
    a, b = o.uniform(n, 1), o.uniform(n, 1, n)
    exact_norm = math.fsum(x * x for x in a.tolist()) if n < 70000 else o.sum_longdouble(a, None, 0)
    assert abs(o.bs3_norm(a) - exact_norm) <= 1e-14 * abs(exact_norm)
    exact_dot = math.fsum((a * b).tolist()) if n < 70000 else o.sum_longdouble(a, b, 1)
    assert abs(o.bs4_dot(a, b) - exact_dot) <= 1e-13 * max(abs(exact_dot), 1e-300) + 1e-300


def test_cg_update_closed_form_and_order():
    n = 70000
    Ap, r = o.uniform(n, 3), o.uniform(n, 3, n)
    v, r2 = o.bs5_update(Ap, 0.37, r)
    want_r = np.array([float(np.float64(x) + np.float64(0.37 * y)) for x, y in zip(r, Ap)])
    assert np.array_equal(r2, want_r)
    # the reduction is the BS3 order applied to the updated r
    assert v == o.bs3_norm(r2)
    v2, x2, r3 = o.bs5_full(0.37, Ap, r, r, Ap)
    assert np.array_equal(x2, r + np.array([0.37 * y for y in Ap]))
    assert np.array_equal(r3, Ap - np.array([0.37 * y for y in r]))
    assert v2 == o.bs3_norm(r3)


# ------------------------------------------------------------ hex ids ----
def brute_hex_ids(nx, ny, nz, p):
    nq = p + 1
    Gx, Gy = nx * p + 1, ny * p + 1
    ids = []
    for ez in range(nz):
        for ey in range(ny):
            for ex in range(nx):
                for k in range(nq):
                    for j in range(nq):
                        for i in range(nq):
                            ids.append(1 + (ex * p + i) + (ey * p + j) * Gx + (ez * p + k) * Gx * Gy)
    return np.array(ids, dtype=np.int64)


@pytest.mark.parametrize("mesh", [(1, 1, 1, 1), (2, 1, 1, 2), (2, 3, 2, 3), (3, 3, 3, 1)])
def test_hex_ids_brute_force(mesh):
    nx, ny, nz, p = mesh
    lex = o.hex_ids(*mesh)
    assert np.array_equal(lex, brute_hex_ids(*mesh))
    assert np.unique(lex).size == (nx * p + 1) * (ny * p + 1) * (nz * p + 1)
    mn = o.hex_ids(*mesh, scheme="min")
    # same grouping: the map lex -> min is a bijection of labels
    pairs = np.unique(np.stack([lex, mn]), axis=1)
    assert pairs.shape[1] == np.unique(lex).size == np.unique(mn).size
    # min label = 1 + first local index of the group
    first = {}
    for l, g in enumerate(lex):
        first.setdefault(g, l)
    assert all(mn[l] == 1 + first[g] for l, g in enumerate(lex))


# ------------------------------------------------------------ ogs setup ----
def brute_ogs(ids):
    """First-appearance numbering by dictionary (one rank, all groups positive or as given)."""
    order, members = [], {}
    for l, v in enumerate(ids):
        if v == 0:
            continue
        a = abs(int(v))
        if a not in members:
            members[a] = []
            order.append(a)
        members[a].append(l)
    pos = [a for a in order if any(ids[l] > 0 for l in members[a])]
    neg = [a for a in order if a not in set(pos)]
    rows = pos + neg
    rs, ci = [0], []
    for a in rows:
        ci += members[a]
        rs.append(len(ci))
    return np.array(rs), np.array(ci), len(pos)


@pytest.mark.parametrize("mesh", [(1, 1, 1, 1), (2, 2, 1, 2), (3, 2, 2, 3)])
@pytest.mark.parametrize("scheme", ["lex", "min"])
def test_ogs_setup_first_appearance(mesh, scheme):
    ids = o.hex_ids(*mesh, scheme=scheme)
    og = o.Ogs(ids, kind=1, unique=True)
    rs, ci, npos = brute_ogs(np.abs(ids))
    assert np.array_equal(og.rowStartsT, rs) and np.array_equal(og.colIdsT, ci)
    assert og.NrowsN == og.NrowsT == npos
    # unique: exactly one positive member per row, N-CSR = the positive members
    flags = og.ids[og.colIdsT] > 0
    per_row = np.add.reduceat(flags.astype(int), og.rowStartsT[:-1])
    assert np.all(per_row == 1)
    assert np.array_equal(og.colIdsN, og.colIdsT[flags])
    # the id scheme does not change the operator
    other = o.Ogs(o.hex_ids(*mesh, scheme="lex" if scheme == "min" else "min"))
    assert np.array_equal(other.rowStartsT, og.rowStartsT) and np.array_equal(other.colIdsT, og.colIdsT)


def test_ogs_setup_signed_flags_and_zeros():
    rng = np.random.default_rng(3)
    for trial in range(30):
        n = int(rng.integers(1, 400))
        ids = rng.integers(-40, 41, size=n).astype(np.int64)
        og = o.Ogs(ids, kind=1, unique=False)
        rs, ci, npos = brute_ogs(ids)
        assert np.array_equal(og.rowStartsT, rs) and np.array_equal(og.colIdsT, ci)
        assert og.NrowsN == npos
        groups = {}
        for v in ids[ids != 0]:
            groups.setdefault(abs(int(v)), []).append(v > 0)
        assert og.gather_defined == all(sum(g) == 1 for g in groups.values())
        un = o.Ogs(ids, kind=0, unique=False)
        assert np.array_equal(un.colIdsN, un.colIdsT)


def test_ogs_setup_rejects_bad_kinds():
    ids = np.arange(1, 10, dtype=np.int64)
    assert o.Ogs(ids, kind=0, unique=True).error == 1
    assert o.Ogs(ids, kind=2, unique=True).error == 1


def multiplicity_counts(nx, ny, nz, p):
    def axis(n):
        interior_faces = n - 1
        total = n * p + 1
        return {1: total - interior_faces, 2: interior_faces}
    out = {}
    for mx, cx in axis(nx).items():
        for my, cy in axis(ny).items():
            for mz, cz in axis(nz).items():
                out[mx * my * mz] = out.get(mx * my * mz, 0) + cx * cy * cz
    return out


@pytest.mark.parametrize("mesh", [(2, 2, 2, 1), (3, 2, 4, 2), (4, 4, 4, 3)])
def test_gather_scatter_multiplicity_kat(mesh):
    ids = o.hex_ids(*mesh)
    og = o.Ogs(ids)
    out = og.gather_scatter(np.ones(ids.size))
    g = og.gather(np.ones(ids.size))
    counts = {int(k): int(v) for k, v in zip(*np.unique(g, return_counts=True))}
    assert counts == multiplicity_counts(*mesh)
    assert set(np.unique(out)) <= {1.0, 2.0, 4.0, 8.0}


def test_gather_invariants_and_composition():
    ids = o.hex_ids(3, 2, 2, 3)
    og = o.Ogs(ids)
    q = np.floor(o.uniform(ids.size, 4) * 1000)  # integers: sums exact
    g = og.gather(q, 1, 0, 2)
    assert g.sum() == q.sum()
    s = og.scatter(g, np.zeros(ids.size), 1, 1)
    assert np.array_equal(s, og.gather_scatter(q))
    qf = o.uniform(ids.size, 4)
    assert np.array_equal(og.scatter(og.gather(qf, 1, 0, 2), np.zeros(ids.size), 1, 1), og.gather_scatter(qf))


def test_order_probe_corner_row():
    ids = o.hex_ids(2, 2, 2, 1)
    og = o.Ogs(ids)
    r = int(np.argmax(np.diff(og.rowStartsT) == 8))
    m = og.colIdsT[og.rowStartsT[r]:og.rowStartsT[r + 1]]
    q = np.zeros(ids.size)
    q[m[0]], q[m[1]], q[m[2]] = 1e16, 1.0, -1e16
    out = og.gather_scatter(q)
    # ascending local order: (1e16 + 1) - 1e16 = 0 since 1e16 + 1 rounds to 1e16
    assert np.all(out[m] == 0.0)
    assert (1e16 - 1e16) + 1.0 == 1.0  # the other order would give 1


@pytest.mark.parametrize("mesh", [(2, 2, 2, 1), (3, 2, 4, 2), (2, 3, 2, 5)])
def test_csr_gather_scatter_pins(mesh):
    """The CSR-array gather and scatter: multiplicity KAT, exact sum invariant on integers, the
    scatter's closed form by numpy indexing (untouched entries included), the composition with the
    gather+scatter, and the ascending-order probe on a corner row."""
    ids = o.hex_ids(*mesh)
    og = o.Ogs(ids)
    rs, ci = og.rowStartsT, og.colIdsT
    g1 = o.csr_gather(rs, ci, np.ones(ids.size))
    counts = {int(k): int(v) for k, v in zip(*np.unique(g1, return_counts=True))}
    assert counts == multiplicity_counts(*mesh)
    q = np.floor(o.uniform(ids.size, 6) * 1000)
    gq = o.csr_gather(rs, ci, q)
    assert gq.sum() == q.sum()
    row_of = np.repeat(np.arange(rs.size - 1), np.diff(rs))
    base = o.uniform(ids.size, 7)
    keep = ci[rs[1]:rs[2]]                       # row 1 is emptied in the scatter's CSR below
    rs_cut = np.concatenate([rs[:2], rs[2:] - (rs[2] - rs[1])]).astype(np.int32)
    ci_cut = np.concatenate([ci[:rs[1]], ci[rs[2]:]]).astype(np.int32)
    gv = o.uniform(rs.size - 1, 8)
    out = o.csr_scatter(rs_cut, ci_cut, gv, base)
    want = base.copy()
    want[ci] = gv[row_of]
    want[keep] = base[keep]
    assert np.array_equal(out, want)
    qf = o.uniform(ids.size, 9)
    assert np.array_equal(o.csr_scatter(rs, ci, o.csr_gather(rs, ci, qf), np.zeros(ids.size)), o.csr_gs(rs, ci, qf))
    assert np.array_equal(o.csr_gather(rs, ci, qf), og.gather(qf, 1, 0, 2))
    r = int(np.argmax(np.diff(rs) == np.diff(rs).max()))
    m = ci[rs[r]:rs[r + 1]]
    probe = np.zeros(ids.size)
    probe[m[0]], probe[m[1]], probe[m[-1]] = 1e16, -1e16, 1.0
    assert o.csr_gather(rs, ci, probe)[r] == 1.0     # (1e16 - 1e16) + 1: the big pair cancels first
    probe[m[1]], probe[m[-1]] = 1.0, -1e16
    assert o.csr_gather(rs, ci, probe)[r] == 0.0     # (1e16 + 1) rounds to 1e16 before - 1e16


def test_okl_block_form_agrees_and_blocks_valid():
    ids = o.hex_ids(8, 8, 8, 7)
    og = o.Ogs(ids)
    bs, rs = og.blockStartsT, og.rowStartsT
    assert bs[0] == 0 and bs[-1] == og.NrowsT and np.all(np.diff(bs) > 0)
    assert np.all(rs[bs[1:]] - rs[bs[:-1]] <= 512)
    q = o.uniform(ids.size, 0)
    assert np.array_equal(o.gs_okl(bs, rs, og.colIdsT, rs, og.colIdsT, q), og.gather_scatter(q))


def test_mul_max_min_ops():
    ids = np.array([1, 2, 1, 3, 2, 1], dtype=np.int64)
    og = o.Ogs(ids, kind=0, unique=False)
    q = np.array([2.0, -3.0, 5.0, 7.0, 4.0, -1.0])
    assert np.array_equal(og.gather_scatter(q, 1, 1, 0), [-10.0, -12.0, -10.0, 7.0, -12.0, -10.0])
    assert np.array_equal(og.gather_scatter(q, 1, 2, 0), [5.0, 4.0, 5.0, 7.0, 4.0, 5.0])
    assert np.array_equal(og.gather_scatter(q, 1, 3, 0), [-1.0, -3.0, -1.0, 7.0, -3.0, -1.0])
    qi = np.array([3, 1, 4, 1, 5, 9], dtype=np.int32)
    assert np.array_equal(og.gather_scatter(qi, 1, 0, 0), [16, 6, 16, 1, 6, 16])


# ---------------------------------------------------------------------------
# reduced-precision storage (FP32 / BF16) restatements
# ---------------------------------------------------------------------------

def test_bf16_rounding_known_answers(oracle):
    f = oracle.lib().oracle_bf16_from_double
    assert f(1.0) == 0x3F80 and f(-2.0) == 0xC000 and f(0.0) == 0 and f(-0.0) == 0x8000
    assert f(1.0 + 2.0**-8) == 0x3F80            # tie -> even (down)
    assert f(1.0 + 3 * 2.0**-8) == 0x3F82        # tie -> even (up)
    assert f(1.0 + 2.0**-8 + 2.0**-40) == 0x3F81  # above the tie: a f64->f32->bf16 chain gets 0x3F80
    assert f(1.0 + 2.0**-8 - 2.0**-40) == 0x3F80
    assert f(float("inf")) == 0x7F80 and f(-float("inf")) == 0xFF80
    assert f(3.0e38) == 0x7F62                   # 3e38 -> 1.1015625 * 2^127 (nearest bf16)
    big = (2 - 2.0**-8) * 2.0**127              # midpoint of the largest bf16 and 2^128: ties to even (inf)
    assert f(big) == 0x7F80 and f(-big) == 0xFF80
    assert f(np.nextafter(big, 0)) == 0x7F7F and f(1e39) == 0x7F80 and f(-1e300) == 0xFF80
    assert f(float("nan")) == 0x7FC0 and f(-float("nan")) == 0x7FC0
    assert f(2.0**-133) == 0x0001 and f(2.0**-134) == 0 and f(1.5 * 2.0**-134) == 0x0001   # denormals


def test_bf16_vectorised_matches_scalar(oracle):
    """The vectorised rounding (large arrays) against the scalar C rule, specials included."""
    f = oracle.lib().oracle_bf16_from_double
    big = (2 - 2.0**-8) * 2.0**127
    v = np.concatenate([oracle.uniform(5000, 8) * np.exp2((oracle.rand_int(5000, 9) % 300) - 150.0),
                        [0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, big, -big, np.nextafter(big, 0), 1e39,
                         -1e300, 2.0**-133, 2.0**-134, 1.5 * 2.0**-134, 3.0e38, 1.0 + 2.0**-8 + 2.0**-40]])
    want = np.array([f(float(x)) for x in v], dtype=np.uint16)
    got = oracle.to_storage(np.concatenate([v] * 2), oracle.STORE_BF16)[: v.size]   # >= 4096: vectorised path
    assert np.array_equal(got, want)


def test_bf16_from_float_matches_torch(oracle):
    torch = pytest.importorskip("torch")
    u = oracle.uniform(20000, 3) * np.exp2((oracle.rand_int(20000, 4) % 60) - 30.0)
    f32 = u.astype(np.float32)
    mine = oracle.to_storage(f32.astype(np.float64), oracle.STORE_BF16)   # f32 values are exact in f64
    ref = torch.from_numpy(f32).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(mine, ref)
    g = oracle.lib().oracle_bf16_from_float
    assert all(g(float(x)) == int(y) for x, y in zip(f32[:500], ref[:500]))


def test_bf16_from_double_brute_force_nearest(oracle):
    # nearest-even by brute force: compare against the two bf16 neighbours in exact arithmetic
    from fractions import Fraction
    u = oracle.uniform(3000, 9) * np.exp2((oracle.rand_int(3000, 10) % 20) - 10.0)
    got = oracle.to_storage(u, oracle.STORE_BF16)
    back = oracle.from_storage(got, oracle.STORE_BF16)
    for x, b, bits in zip(u, back, got):
        e = int(np.frexp(abs(x))[1])             # |x| in [2^(e-1), 2^e)
        ulp = Fraction(2) ** (e - 1 - 7)
        fx, fb = Fraction(float(x)), Fraction(float(b))
        assert abs(fx - fb) <= ulp / 2
        if abs(fx - fb) == ulp / 2:
            assert int(bits) % 2 == 0


def test_gs_mixed_f64_storage_is_plain_gather_scatter(oracle):
    ids = oracle.hex_ids(3, 2, 2, 3)
    o = oracle.Ogs(ids)
    q = oracle.uniform(ids.size, 1)
    assert np.array_equal(oracle.gs_mixed(oracle.STORE_F64, o.rowStartsT, o.colIdsT, q), o.gather_scatter(q))


def test_gs_mixed_rounds_once(oracle):
    # members 1, 2^-24, 2^-24 in f32: a float fold drops both small terms (ties to even),
    # one rounding of the f64 sum keeps them
    rs = np.array([0, 3], dtype=np.int32)
    ci = np.array([0, 1, 2], dtype=np.int32)
    q = np.array([1.0, 2.0**-24, 2.0**-24], dtype=np.float32)
    out = oracle.gs_mixed(oracle.STORE_F32, rs, ci, q)
    assert np.all(out == np.float32(1.0 + 2.0**-23))
    # bf16 (ulp 2^-7 at 1): 1 + 2 * 3*2^-10 rounds up; a bf16 fold loses each 3*2^-10 term
    qb = oracle.to_storage(np.array([1.0, 3 * 2.0**-10, 3 * 2.0**-10]), oracle.STORE_BF16)
    outb = oracle.from_storage(oracle.gs_mixed(oracle.STORE_BF16, rs, ci, qb), oracle.STORE_BF16)
    assert np.all(outb == 1.0 + 2.0**-7)


def test_gs_mixed_error_bound_vs_f64(oracle):
    ids = oracle.hex_ids(4, 3, 3, 5)
    o = oracle.Ogs(ids)
    q = oracle.uniform(ids.size, 2)
    exact = o.gather_scatter(q)
    for st, eps in ((oracle.STORE_F32, 2.0**-24), (oracle.STORE_BF16, 2.0**-8)):
        qs = oracle.to_storage(q, st)
        got = oracle.from_storage(oracle.gs_mixed(st, o.rowStartsT, o.colIdsT, qs), st)
        # input rounding (<= eps |q_i| per member) plus one output rounding (<= eps |sum|)
        absq = np.abs(q)
        bound = eps * o.gather_scatter(absq) + eps * np.abs(exact) + 1e-300
        assert np.all(np.abs(got - exact) <= bound * (1 + 1e-12))


def test_cg_mixed_closed_form(oracle):
    # dyadic values: every FP32 operation exact, so the result is the real-number update
    n = 1001
    p = ((oracle.rand_int(n, 1) % 64) - 32).astype(np.float32) / 8
    Ap = ((oracle.rand_int(n, 2) % 64) - 32).astype(np.float32) / 16
    x = ((oracle.rand_int(n, 3) % 64) - 32).astype(np.float32) / 4
    r = ((oracle.rand_int(n, 4) % 64) - 32).astype(np.float32) / 2
    v, xo, ro = oracle.cg_mixed(oracle.STORE_F32, 0.5, p, Ap, x, r)
    xe = x.astype(np.float64) + 0.5 * p
    re = r.astype(np.float64) - 0.5 * Ap
    assert np.array_equal(xo, xe.astype(np.float32)) and np.array_equal(ro, re.astype(np.float32))
    assert v == float(np.sum(re * re))


def test_cg_mixed_no_fma_and_store_rounding(oracle):
    # r - a*Ap with one rounding per operation: a*Ap = (1+2^-12)^2 rounds to 1+2^-11 in f32
    a, ap = np.float32(1 + 2.0**-12), np.float32(1 + 2.0**-12)
    p = np.zeros(1, np.float32)
    v, _, ro = oracle.cg_mixed(oracle.STORE_F32, float(a), p, np.array([ap]), p.copy(), np.array([np.float32(1 + 2.0**-11)]))
    assert ro[0] == 0.0 and v == 0.0   # a fused a*Ap would leave -2^-24
    qb = oracle.to_storage(np.array([1.0]), oracle.STORE_BF16)
    zb = oracle.to_storage(np.array([0.0]), oracle.STORE_BF16)
    rb = oracle.to_storage(np.array([2.0**-9]), oracle.STORE_BF16)
    v, xb, rb2 = oracle.cg_mixed(oracle.STORE_BF16, 1.0 + 2.0**-8, zb, zb, qb, rb)
    assert oracle.from_storage(xb, oracle.STORE_BF16)[0] == 1.0   # x unchanged
    assert oracle.from_storage(rb2, oracle.STORE_BF16)[0] == 2.0**-9 and v == 2.0**-18
