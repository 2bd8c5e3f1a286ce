"""Attribute recording on the device (ftk2h_result_add_attributes) against the
oracle's restatement of engine.hpp:1144-1201, bit for bit in float32, for CP2D,
CP3D (with degeneracies, on an offset slab), contour, and batched results; host and
device views."""
import sys

import numpy as np
import pytest

import oracle
import refcases
from ftk2_amd import engine, synth
from ftk2_amd import _lib

pytestmark = pytest.mark.gpu

KIND = {"cp": oracle.PRED_CP, "contour": oracle.PRED_CONTOUR, "fiber": oracle.PRED_FIBER}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.transpose())).cuda().permute(*reversed(range(a.ndim)))


def check(name, comps, kind, m, specs, thr=(0.0, 0.0), offset=None, gdims=None, device_views=False):
    say(f"[attr] {name}: shape={comps[0].shape} specs={[(s, len(a)) for s, a in specs]}")
    shape = comps[0].shape
    off = offset or (0,) * len(shape)
    gd = gdims or shape
    ref = oracle.run(comps, KIND[kind], m, thr=thr, offset=offset, gdims=gdims)
    mesh = engine.RegularSimplicialMesh(shape, offset, gdims)
    res = engine.extract(mesh, KIND[kind], m, comps, thr)
    dspecs = [(s, [to_dev(a) for a in arrs]) for s, arrs in specs] if device_views else specs
    res.add_attributes(mesh, dspecs)
    got = res.attributes()
    want = oracle.attributes(ref["verts"], ref["bary"], specs, gd, off, comps[0].dtype)
    assert got.shape == want.shape
    assert np.array_equal(got, want), f"max diff {np.abs(got - want).max()}"
    return got


def test_cp2d_attributes():
    u, v = synth.cp2d_config(nx=48, ny=40, nt=6, ngauss=8, seed=4)
    s = synth.gaussians((48, 40), 6, 6, seed=9, gradient=False)[0]
    check("cp2d", [u, v], "cp", 2, [(0, [s]), (1, [u, v]), (15, [v]), (1, [s])])


def test_cp2d_attributes_device_views():
    u, v = synth.cp2d_config(nx=33, ny=41, nt=4, ngauss=8, seed=5)
    check("cp2d-dev", [u, v], "cp", 2, [(4, [u, v]), (6, [u])], device_views=True)


def test_cp3d_attributes_offset_slab():
    comps = synth.cp3d_config(n=10, nt=3, ngauss=4, seed=2, zero_frac=0.03, dup_frac=0.03)
    s = synth.gaussians((10, 10, 10), 3, 4, seed=3, gradient=False)[0]
    check("cp3d-slab", comps, "cp", 3, [(0, [s]), (2, comps)], offset=(0, 0, 0, 2), gdims=(10, 10, 10, 6))


def test_contour_and_float_attributes():
    f = refcases.levelset3d_float()[0][:20, :18, :6]
    f = np.asfortranarray(f)
    check("contour-float", [f], "contour", 1, [(3, [f]), (8, [f, f])])


def test_batched_results_attributes():
    jobs = []
    for i in range(4):
        u, v = synth.cp2d_config(nx=20 + 5 * i, ny=24, nt=4, ngauss=6, seed=30 + i)
        jobs.append((engine.RegularSimplicialMesh(list(u.shape)), [u, v]))
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    res = b.run(jobs)
    for (mesh, comps), r in zip(jobs, res):
        r.add_attributes(mesh, [(1, [comps[0], comps[1]])])
        ref = oracle.run(comps, oracle.PRED_CP, 2)
        want = oracle.attributes(ref["verts"], ref["bary"], [(1, comps)], comps[0].shape, (0, 0, 0), np.float64)
        assert np.array_equal(r.attributes(), want)
