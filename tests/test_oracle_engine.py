"""Oracle pins end to end: golden counts held by the reference's own tests
(tests/test_examples.cpp, test_engine.cpp, test_mesh.cpp) and invariants."""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.csgraph

import oracle
import refcases


def test_levelset3d_golden_66418():  # test_examples.cpp:229-257
    r = oracle.run(refcases.levelset3d_float(), oracle.PRED_CONTOUR, 1)
    assert len(r["verts"]) == 66418


def test_fiber3d_golden_6078():  # test_examples.cpp:188-227
    r = oracle.run(refcases.fiber3d_double(), oracle.PRED_FIBER, 2)
    assert len(r["verts"]) == 6078


def test_cp2d_woven_golden_788():  # test_examples.cpp:39-105
    u, v, s = refcases.woven_cp2d()
    r = oracle.run([u, v], oracle.PRED_CP, 2, dims=(14, 14, 5), offset=(1, 1, 0), gdims=(16, 16, 5), scalar=s)
    assert len(r["verts"]) == 788
    # the time parametrisation matters: integer times give another complex
    tt = np.arange(5, dtype=np.float64)[None, None, :]
    x = (np.arange(16)[:, None, None] / 15 - 0.5) * 15.0
    y = (np.arange(16)[None, :, None] / 15 - 0.5) * 15.0
    s2 = np.asfortranarray(np.cos(x * np.cos(tt) - y * np.sin(tt)) * np.sin(x * np.sin(tt) + y * np.cos(tt)))
    u2 = np.zeros_like(s2, order="F")
    v2 = np.zeros_like(s2, order="F")
    u2[1:-1, 1:-1, :] = (s2[2:, 1:-1, :] - s2[:-2, 1:-1, :]) / 2.0
    v2[1:-1, 1:-1, :] = (s2[1:-1, 2:, :] - s2[1:-1, :-2, :]) / 2.0
    r2 = oracle.run([u2, v2], oracle.PRED_CP, 2, dims=(14, 14, 5), offset=(1, 1, 0), gdims=(16, 16, 5))
    assert len(r2["verts"]) != 788


def test_cp3d_linear_float():
    """test_examples.cpp:153-186 asserts 52 vertices.  A literal transcription of the
    reference sources (sos.hpp det3 with its explicit 4x4 cofactor formula,
    zero_crossing.hpp, the all-faces node set of engine.hpp) yields 336 SoS-inside
    tetrahedra of which 29 pass the solver gate; the oracle agrees with the
    transcription.  Pinned here: 29 nodes, one track."""
    r = oracle.run(refcases.cp3d_linear_float(), oracle.PRED_CP, 3)
    assert len(r["verts"]) == 29
    assert r["n_components"] == 1


def test_engine_cp1d_structure():  # test_engine.cpp:36-51
    r = oracle.run(refcases.cp1d_engine(), oracle.PRED_CP, 1)
    assert len(r["verts"]) > 0
    assert 1 in r["conn"]


def test_engine_levelset2d_structure():  # test_engine.cpp:53-71
    r = oracle.run(refcases.levelset2d_engine(), oracle.PRED_CONTOUR, 1)
    assert len(r["verts"]) > 0
    assert 2 in r["conn"]


def _components(n, cells):
    if len(cells) == 0:
        return np.arange(n)
    rows = np.repeat(cells[:, 0], cells.shape[1] - 1)
    cols = cells[:, 1:].ravel()
    g = scipy.sparse.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
    _, lab = scipy.sparse.csgraph.connected_components(g, directed=False)
    # track id = minimum index in the component
    mins = np.full(lab.max() + 1, n)
    np.minimum.at(mins, lab, np.arange(n))
    return mins[lab]


@pytest.mark.parametrize("case", ["levelset", "fiber", "cp2d"])
def test_track_ids_are_component_minima(case):
    if case == "levelset":
        r = oracle.run(refcases.levelset3d_float(), oracle.PRED_CONTOUR, 1)
        cells = r["conn"][3]
    elif case == "fiber":
        r = oracle.run(refcases.fiber3d_double(), oracle.PRED_FIBER, 2)
        cells = r["conn"][2]
    else:
        from ftk2_amd import synth
        u, v = synth.cp2d_config(seed=3)
        r = oracle.run([u, v], oracle.PRED_CP, 2)
        cells = r["conn"][1]
    ref = _components(len(r["verts"]), cells.astype(np.int64))
    assert np.array_equal(r["track"].astype(np.int64), ref)


def test_connectivity_sorted_unique():
    r = oracle.run(refcases.levelset3d_float(), oracle.PRED_CONTOUR, 1)
    c = r["conn"][3]
    assert np.all(np.diff(c, axis=1) > 0)
    keys = [tuple(x) for x in c.tolist()]
    assert keys == sorted(set(keys))


def test_node_set_is_every_face_of_every_cell():
    """Brute force on a small CP2D case: the node set equals every triangle of every
    Kuhn cell that passes the predicate (LUT pass + upper-boundary faces)."""
    from ftk2_amd import synth
    u, v = synth.gaussians((9, 8), 4, 4, seed=11)
    r = oracle.run([u, v], oracle.PRED_CP, 2)
    got = {tuple(x) for x in r["verts"].tolist()}
    nx, ny, nt = u.shape
    gid = lambda x, y, t: x + nx * (y + ny * t)
    import itertools
    tris = set()
    for t in range(nt - 1):
        for y in range(ny - 1):
            for x in range(nx - 1):
                for perm in itertools.permutations(range(3)):
                    c = [0, 0, 0]
                    vs = [(x, y, t)]
                    for ax in perm:
                        c[ax] += 1
                        vs.append((x + c[0], y + c[1], t + c[2]))
                    for tri in itertools.combinations(vs, 3):
                        tris.add(tuple(sorted(gid(*p) for p in tri)))
    inside = set()
    for tri in tris:
        pts = []
        for g in tri:
            x, r_ = g % nx, g // nx
            y, t = r_ % ny, r_ // ny
            pts.append((x, y, t))
        vals = [[u[p], v[p]] for p in pts]
        if oracle.origin_inside(vals, list(tri)) and oracle.solve(vals)[0]:
            inside.add(tri)
    assert got == inside


def test_stream_equals_oneshot():  # test_examples.cpp:259-297 contract
    from ftk2_amd import synth
    u, v = synth.gaussians((20, 18), 5, 6, seed=5)
    one = oracle.run([u, v], oracle.PRED_CP, 2)
    steps = [[np.asfortranarray(u[:, :, t]), np.asfortranarray(v[:, :, t])] for t in range(5)]
    st = oracle.run_stream(steps, oracle.PRED_CP, 2)
    assert np.array_equal(one["verts"], st["verts"])
    assert np.array_equal(one["track"], st["track"])
    assert np.array_equal(one["conn"][1], st["conn"][1])


@pytest.mark.parametrize("case", ["fiber", "contour", "cp3d", "cp3d_float"])
def test_stream_equals_oneshot_every_predicate(case):
    """execute_stream (one [t, t+1] slab per step, engine.hpp:214-270) equals execute()
    on the whole grid for every predicate, lambda bits and all cell dimensions
    included: the stream restatement the GPU stream tests compare with."""
    from ftk2_amd import synth
    if case == "fiber":
        comps = [synth.scalar_config(n=14, nt=5, ngauss=8, seed=s, dtype=np.float64) for s in (7, 8)]
        kind, m, thr = oracle.PRED_FIBER, 2, (0.0, 0.0)
    elif case == "contour":
        comps = [synth.scalar_config(n=16, nt=6, ngauss=8, seed=1, dtype=np.float64)]
        kind, m, thr = oracle.PRED_CONTOUR, 1, (0.0, 0.0)
    elif case == "cp3d":
        comps, kind, m, thr = synth.cp3d_config(n=14, nt=5, seed=3, noise=0.02), oracle.PRED_CP, 3, (0.0, 0.0)
    else:
        comps = synth.gaussians((14, 13, 12), 5, 6, sigma=(8, 24), seed=4, noise=0.02, zero_frac=0.02,
                                dup_frac=0.02, dtype=np.float32)
        kind, m, thr = oracle.PRED_CP, 3, (0.0, 0.0)
    one = oracle.run(comps, kind, m, thr=thr, nthreads=4)
    steps = [[np.asfortranarray(c[..., t]) for c in comps] for t in range(comps[0].shape[-1])]
    st = oracle.run_stream(steps, kind, m, thr=thr, nthreads=4)
    assert len(one["verts"]) > 0
    for k in ("verts", "track", "bary", "scalar"):
        assert np.array_equal(one[k], st[k]), k
    assert np.array_equal(one["lam"].view(np.uint64), st["lam"].view(np.uint64))
    assert sorted(one["conn"]) == sorted(st["conn"])
    for d in one["conn"]:
        assert np.array_equal(one["conn"][d], st["conn"][d])


def test_field_arrays_must_share_one_dtype():
    """The C side reads every component with the first one's element size, so the
    binding refuses components of mixed dtypes (a float32 array read as float64
    would be read past its end) and dtypes other than float32 / float64."""
    a = np.zeros((6, 5, 3), np.float32)
    b = np.zeros((6, 5, 3), np.float64)
    with pytest.raises(TypeError):
        oracle.run([a, b], oracle.PRED_FIBER, 2)
    steps = [[np.asfortranarray(a[..., t]), np.asfortranarray(b[..., t])] for t in range(3)]
    with pytest.raises(TypeError):
        oracle.run_stream(steps, oracle.PRED_FIBER, 2)
    with pytest.raises(TypeError):
        oracle.run([np.zeros((6, 5, 3), np.int32)], oracle.PRED_CONTOUR, 1)
    oracle.run([a, a.copy()], oracle.PRED_FIBER, 2)  # one dtype: accepted
