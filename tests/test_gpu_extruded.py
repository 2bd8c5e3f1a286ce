"""ExtrudedSimplicialMesh over a regular base on the GPU (ftk2h_mesh_desc
time_id_stride): nodes (extruded vertex ids), lambda, scalar, tracks and cells
equal the oracle run with the same id scheme, bit for bit; contour sign ties take
the parity of the extruded ids; device positions carry the layer as Time."""
import sys

import numpy as np
import pytest

import oracle
from ftk2_amd import engine, synth, vtk

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare(name, comps, kind, m, thr=(0.0, 0.0)):
    spatial = list(comps[0].shape[:-1])
    nt = comps[0].shape[-1]
    mesh = engine.ExtrudedSimplicialMesh(engine.RegularSimplicialMesh(spatial), nt - 1)
    say(f"[extruded] {name} spatial={spatial} layers={nt - 1} N={mesh.time_stride}")
    res = engine.extract(mesh, kind, m, comps, thresholds=thr)
    cx = res.to_complex()
    okind = {0: oracle.PRED_CP, 1: oracle.PRED_CONTOUR, 2: oracle.PRED_FIBER}[kind]
    ref = oracle.run(comps, okind, m, thr=thr, time_stride=mesh.time_stride)
    assert len(ref["verts"]) > 0
    assert np.array_equal(cx.vertices, ref["verts"].astype(np.uint64)), "node simplices differ"
    assert np.array_equal(cx.lam, ref["lam"]), "lambda differs"
    assert np.array_equal(cx.scalar, ref["scalar"])
    assert np.array_equal(cx.track_id, ref["track"].astype(np.uint64))
    assert sorted(cx.connectivity) == sorted(ref["conn"])
    for d in ref["conn"]:
        assert np.array_equal(cx.connectivity[d], ref["conn"][d])
    dv = res.device_view()  # what a consumer needs to decode keys on the device
    assert dv.time_id_stride == mesh.time_stride and dv.n_types == len(oracle.chains(len(mesh.dims), m))
    assert list(dv.global_dims)[:len(mesh.dims)] == mesh.global_dims
    p, t = res.positions(with_time=True)
    ep, et = vtk.positions(cx.vertices, cx.bary, mesh.global_dims, mesh.time_stride, with_time=True)
    assert np.array_equal(p, ep) and np.array_equal(t, et)
    op, ot = oracle.positions(cx.vertices, cx.bary, mesh.global_dims, mesh.time_stride)  # utils/vtk.hpp:26-59
    assert np.array_equal(p, op) and np.array_equal(t, ot)
    return cx


def test_extruded_cp2d():
    u, v = synth.gaussians((23, 19), 6, 6, seed=31, zero_frac=0.02, dup_frac=0.02)
    compare("cp2d", [u, v], 0, 2)


def test_extruded_cp3d():
    c = synth.cp3d_config(n=12, nt=4, ngauss=4, seed=32, noise=0.02, zero_frac=0.02, dup_frac=0.02)
    compare("cp3d", c, 0, 3)


def test_extruded_fiber():
    a = synth.gaussians((14, 12, 10), 3, 5, gradient=False, seed=33, subtract_mean=True, dtype=np.float32)[0]
    b = synth.gaussians((14, 12, 10), 3, 5, gradient=False, seed=34, subtract_mean=True, dtype=np.float32)[0]
    compare("fiber", [a, b], 2, 2, (0.001, -0.002))


def test_extruded_contour_ties():
    nx, ny, nt = 9, 6, 4  # N = 9 + 54 = 63 (odd) vs 54: tie parities differ at odd t
    rng = np.random.default_rng(35)
    f = np.asfortranarray(rng.uniform(-1, 1, size=(nx, ny, nt)))
    for p in [(1, 1, 1), (3, 2, 1), (5, 4, 3), (2, 3, 2), (7, 1, 3), (4, 4, 1)]:
        f[p] = 0.25
    cx = compare("contour", [f], 1, 1, (0.25, 0.0))
    reg = oracle.run([f], oracle.PRED_CONTOUR, 1, thr=(0.25, 0.0))
    assert len(cx.vertices) != len(reg["verts"]) or not np.array_equal(
        cx.vertices % np.uint64(63) + (cx.vertices // np.uint64(63)) * np.uint64(54), reg["verts"].astype(np.uint64))
