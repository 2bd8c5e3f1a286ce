"""The oracle's full-size mode (oracle/lean.hpp) equals its engine restatement
(oracle.run, engine.hpp execute()/get_complex()) on small meshes, before it is
trusted on the 256^3 x 4 critical-point workload: node vertices, lambda bits, track
ids and star edges, with zero and duplicated vertices, offset meshes, FP32 and FP64,
and the fiber predicate on 2-D + time."""
import numpy as np
import pytest

import oracle
from ftk2_amd import synth


def same(comps, kind, m, **kw):
    ref = oracle.run(comps, kind, m, nthreads=4, **kw)
    L = oracle.Lean(comps, kind, m, nthreads=4, **kw)
    verts, lam, track = L.nodes()
    edges = L.edges()
    assert L.n_nodes == len(ref["verts"]) > 0
    assert np.array_equal(verts, ref["verts"])
    assert np.array_equal(lam.view(np.uint64), ref["lam"].view(np.uint64))
    assert np.array_equal(track, ref["track"])
    assert np.array_equal(edges, ref["conn"][1])
    assert L.n_components == ref["n_components"]
    # ranges read the same rows
    v2, l2, t2 = L.nodes(3, 5)
    assert np.array_equal(v2, verts[3:8]) and np.array_equal(t2, track[3:8])
    assert np.array_equal(L.edges(2, 4), edges[2:6])
    L.close()
    return ref


@pytest.mark.parametrize("seed", [0, 1, 3])
def test_cp3d_degenerate(seed):
    comps = synth.cp3d_config(n=(16, 14, 13), nt=4, seed=seed, noise=0.05)
    ref = same(comps, oracle.PRED_CP, 3)
    assert len(ref["conn"][1]) > 0


def test_cp3d_float():
    comps = [c.astype(np.float32, order="F") for c in synth.cp3d_config(n=(9, 10, 8), nt=3, seed=5)]
    same(comps, oracle.PRED_CP, 3)


def test_cp2d_offset_mesh():
    u, v = synth.cp2d_config(nx=40, ny=36, nt=6, seed=3)
    sub = [np.asfortranarray(a[2:38, 1:35, 1:6]) for a in (u, v)]
    same(sub, oracle.PRED_CP, 2, offset=(2, 1, 1), gdims=(40, 36, 6))


def test_cp2d_config0():
    same(list(synth.cp2d_config(seed=0)), oracle.PRED_CP, 2)


def test_fiber_2d():
    f = synth.gaussians((30, 28), 5, 12, seed=9, gradient=False, subtract_mean=True)[0]
    g = synth.gaussians((30, 28), 5, 12, seed=10, gradient=False, subtract_mean=True)[0]
    same([f, g], oracle.PRED_FIBER, 2, thr=(0.01, -0.02))


def test_rejects_other_cells():
    f = synth.gaussians((8, 8, 8), 3, 4, seed=1, gradient=False)[0]
    with pytest.raises(RuntimeError):
        oracle.Lean([f], oracle.PRED_CONTOUR, 1)
