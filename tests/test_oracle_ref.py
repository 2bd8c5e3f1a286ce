"""The oracle's restatement against the reference's own code: sos.hpp and mesh.hpp
(hguo/ftk2 include/ftk2/core/) compiled from the reference tree by
`make -C oracle ref` (oracle/ref_driver.cpp, oracle/ref.py).  Skipped where the
reference tree is absent.

Pinned here: sign / det2 / det3 (all fifteen det3 levels, through exhaustive and
random degenerate integer cases) and origin_inside<1..3> (sos.hpp:12-254); the simplex
lookup tables and the LUT enumeration (mesh.hpp:111-302, which misses the
upper-boundary simplices); encode_simplex_id (mesh.hpp:736-749); and that the d = 4
cofaces pass of get_complex (engine.hpp:896-963, Mesh::cofaces / faces) unites no two
nodes that the star edges left in different tracks."""
import itertools

import numpy as np
import pytest

import oracle
from oracle import ref

pytestmark = pytest.mark.skipif(not ref.available() and not ref.build(), reason="reference tree not present")


def both_inside(vals, idx, dtype=1):
    a = oracle.origin_inside_many(vals, idx, dtype=dtype)
    b = ref.origin_inside_many(vals, idx, dtype=dtype)
    return a, b


def test_origin_inside3_exhaustive_unit_box():
    """Every tetrahedron with coordinates in {-1, 0, 1} / Q: 3^12 cases, deep into the
    fallback levels of det3 (zero, repeated and collinear vertices).  A sign flipped in
    any one det3 level of the oracle fails this test, except level 9 (rows v0, v2,
    origin; components 1, 2), which is never reached: levels 5, 6 and 8 are the three
    components of v2, so they are all zero only when v2 is, and then so is level 9."""
    grid = np.array(list(itertools.product((-1.0, 0.0, 1.0), repeat=12))) * 1e-6
    vals = grid.reshape(-1, 4, 3)
    idx = np.tile(np.array([10, 11, 13, 17], np.uint64), (len(vals), 1))
    a, b = both_inside(vals, idx)
    assert np.array_equal(a, b)
    assert 0 < a.sum() < len(a)


@pytest.mark.parametrize("seed", [0, 1])
def test_origin_inside3_random_degenerate(seed):
    """Small integers (many ties), duplicated and zero vertices, large magnitudes
    (beyond 2^20 and 2^26 after quantization), float and double."""
    rng = np.random.default_rng(seed)
    n = 300000
    vals = rng.integers(-3, 4, size=(n, 4, 3)).astype(np.float64) * 1e-6
    dup = rng.random(n) < 0.2
    vals[dup, 1] = vals[dup, 0]
    zero = rng.random(n) < 0.1
    vals[zero, 2] = 0.0
    big = rng.random(n) < 0.1
    vals[big] *= rng.choice([3e3, 7e7, 1e9], size=(big.sum(), 1, 1))
    cont = rng.random(n) < 0.3
    vals[cont] = rng.normal(size=(cont.sum(), 4, 3))
    idx = np.sort(rng.integers(0, 1 << 40, size=(n, 4)).astype(np.uint64), axis=1)
    for dtype in (1, 0):
        a, b = both_inside(vals, idx, dtype)
        assert np.array_equal(a, b), f"dtype {dtype}: {np.flatnonzero(a != b)[:5]}"


def test_all_zero_simplices_have_a_constant_answer():
    """A simplex whose quantized values are all zero, vertices in ascending id order
    (as every chain simplex of the mesh lists them): origin_inside<3> is always true
    and origin_inside<2> always false, whatever the ids (every det level is zero and
    the last, constant one decides).  The 3-D fast path relies on it."""
    rng = np.random.default_rng(17)
    n = 20000
    idx = np.sort(rng.integers(0, 1 << 40, size=(n, 4)).astype(np.uint64), axis=1)
    keep = np.all(np.diff(idx.astype(np.int64), axis=1) > 0, axis=1)
    idx = idx[keep]
    for dtype in (1, 0):
        a, b = both_inside(np.zeros((len(idx), 4, 3)), idx, dtype)
        assert a.all() and b.all()
        # values below half a quantum (|v| Q < 1/2) round to zero too
        tiny = rng.uniform(-0.49e-6, 0.49e-6, size=(len(idx), 4, 3))
        a, b = both_inside(tiny, idx, dtype)
        assert a.all() and b.all()
        a, b = both_inside(np.zeros((len(idx), 3, 2)), np.ascontiguousarray(idx[:, :3]), dtype)
        assert not a.any() and not b.any()


def test_det3_levels_direct():
    rng = np.random.default_rng(7)
    for _ in range(20000):
        v = rng.integers(-2, 3, size=(3, 3)).astype(np.float64) * 1e-6
        if rng.random() < 0.3:
            v[rng.integers(3)] = v[rng.integers(3)]
        assert oracle.det3(*v) == ref.det3(*v)


def test_origin_inside2_and_1_exhaustive():
    grid = np.array(list(itertools.product((-2.0, -1.0, 0.0, 1.0, 2.0), repeat=6))) * 1e-6
    vals = grid.reshape(-1, 3, 2)
    for ids in ([3, 4, 9], [4, 3, 9], [9, 4, 3], [5, 5, 6]):
        idx = np.tile(np.array(ids, np.uint64), (len(vals), 1))
        a, b = both_inside(vals, idx)
        assert np.array_equal(a, b)
    v1 = np.array(list(itertools.product((-1.0, 0.0, 1.0, 2.5), repeat=2))).reshape(-1, 2, 1) * 1e-6
    for ids in ([2, 3], [3, 2], [4, 6]):
        idx = np.tile(np.array(ids, np.uint64), (len(v1), 1))
        a, b = both_inside(v1, idx)
        assert np.array_equal(a, b)


def test_sign_and_det2_ties():
    for v in (0.0, 4e-7, -4e-7, 5e-7, -5e-7, 1e-6, -3.2):
        for i in range(4):
            assert oracle.sign(v, i) == ref.sign(v, i)
            assert oracle.sign(v, i, dtype=0) == ref.sign(v, i, dtype=0)
    rng = np.random.default_rng(3)
    for _ in range(5000):
        a, b = rng.integers(-2, 3, size=(2, 2)) * 1e-6
        i0, i1 = rng.integers(0, 4, size=2)
        assert oracle.det2(a, b, i0, i1) == ref.det2(a, b, i0, i1)


@pytest.mark.parametrize("d,m", [(2, 1), (2, 2), (3, 1), (3, 2), (3, 3), (4, 1), (4, 2), (4, 3), (4, 4)])
def test_lut_rows_are_the_chains(d, m):
    rows = ref.lut(d, m)
    mine = {tuple(int(x) for x in r) for r in oracle.chains(d, m)}
    theirs = set()
    for r in rows:
        masks = sorted(int(sum(int(r[vi][j]) << j for j in range(d))) for vi in range(m + 1))
        theirs.add(tuple(masks))
    assert len(rows) == len(mine) == len(theirs)
    assert {tuple(sorted(t)) for t in mine} == theirs


@pytest.mark.parametrize("dims,k", [((4, 3), 1), ((4, 3), 2), ((3, 4, 3), 2), ((3, 3, 3), 3), ((3, 2, 3, 3), 3),
                                    ((2, 3, 2, 3), 2), ((3, 3, 2, 2), 1)])
def test_lut_enumeration(dims, k):
    """iterate_simplices visits the chain simplices of every hypercube with its base in
    the grid, each once, and no upper-boundary simplex."""
    d = len(dims)
    got = ref.iterate_simplices(dims, k)
    assert len({tuple(r) for r in got}) == len(got)
    gs = [int(np.prod(dims[:j])) for j in range(d)]
    want = set()
    for base in itertools.product(*[range(n - 1) for n in dims]):
        b = sum(x * g for x, g in zip(base, gs))
        for ch in oracle.chains(d, k):
            want.add(tuple(sorted(b + sum(gs[j] for j in range(d) if (int(mk) >> j) & 1) for mk in ch)))
    assert {tuple(int(x) for x in r) for r in got} == want


def test_encode_simplex_id():
    rng = np.random.default_rng(11)
    for _ in range(3000):
        d = int(rng.integers(2, 5))
        g = [int(x) for x in rng.integers(2, 40, size=d)]
        m = int(rng.integers(1, d + 1))
        ch = oracle.chains(d, m)[rng.integers(len(oracle.chains(d, m)))]
        base = [int(rng.integers(0, n - 1)) for n in g]
        gs = [int(np.prod(g[:j])) for j in range(d)]
        b = sum(x * s for x, s in zip(base, gs))
        verts = [b + sum(gs[j] for j in range(d) if (int(mk) >> j) & 1) for mk in ch]
        rng.shuffle(verts)
        assert oracle.encode_simplex_id(g, verts) == ref.encode_simplex_id(g, verts)


@pytest.mark.parametrize("seed", [0, 5])
def test_cofaces_pass_adds_no_union(seed):
    """For every node tetrahedron, every tet of every pentatope the reference's
    Mesh::cofaces returns for it (then Mesh::faces) that is also a node already has
    the same track id in the oracle's complex: the pass unites nothing, so skipping
    it does not change track_id."""
    from ftk2_amd import synth
    comps = synth.cp3d_config(n=(14, 13, 12), nt=4, seed=seed, noise=0.05)
    r = oracle.run(comps, oracle.PRED_CP, 3)
    verts, track = r["verts"], r["track"]
    assert len(verts) > 50 and len(np.unique(track)) > 1
    index = {tuple(int(x) for x in v): i for i, v in enumerate(verts)}
    dims = comps[0].shape
    checked = 0
    for i, v in enumerate(verts):
        for nb in ref.coface_neighbors(dims, v):
            j = index.get(tuple(int(x) for x in nb))
            if j is not None:
                assert track[j] == track[i], (v, nb)
                checked += 1
    assert checked > len(verts)


def test_golden52_inside_set_from_the_reference_sos():
    """The CP3D float case of test_examples.cpp:153-186 (24^3 x 8, golden 52).  Every
    tetrahedron of the mesh (the reference's own LUT chains from every vertex, all four
    corners in the grid: every face of every Kuhn cell) goes through
    the reference's compiled origin_inside<3> and the oracle's: both accept the same
    336, and the oracle's solver gate keeps 29 of them, the node count the oracle's
    engine gives for this case (test_oracle_engine.py::test_cp3d_linear_float).  Only
    29 of the 336 have corner values of both signs (or zero) in every component: with
    a zero component the det3 levels decide by minors that need not bound the origin,
    so no geometric prefilter is used."""
    import refcases
    comps = refcases.cp3d_linear_float()
    dims = np.array(comps[0].shape, dtype=np.int64)
    vals = np.stack([c.ravel(order="F").astype(np.float64) for c in comps], axis=1)  # x fastest
    base = np.stack(np.meshgrid(*[np.arange(n) for n in dims], indexing="ij"), axis=-1).reshape(-1, 4, order="F")
    stride = np.array([1, dims[0], dims[0] * dims[1], dims[0] * dims[1] * dims[2]], dtype=np.int64)
    L = ref.lut(4, 3)
    assert L.shape == (60, 4, 4)
    n_tets, inside_v = 0, []
    for t in range(len(L)):
        corners = base[:, None, :] + L[t][None, :, :]  # (n, 4 vertices, 4 axes)
        ok = np.all(corners < dims, axis=(1, 2))
        ids = (corners[ok] * stride).sum(axis=2)
        assert np.all(np.diff(ids, axis=1) > 0)  # ascending vertex ids
        v = vals[ids]  # (n, 4, 3)
        a = ref.origin_inside_many(v, ids.astype(np.uint64), dtype=0)
        b = oracle.origin_inside_many(v, ids.astype(np.uint64), dtype=0)
        assert np.array_equal(a, b), t
        n_tets += len(ids)
        inside_v.append(v[a])
    inside_v = np.concatenate(inside_v)
    # a type whose top corner raises axis j has dims_j - 1 positions along j, else dims_j
    assert n_tets == sum(int(np.prod(dims - L[t][3])) for t in range(len(L)))
    assert len(inside_v) == 336
    assert sum(oracle.solve(x, dtype=0)[0] for x in inside_v) == 29
