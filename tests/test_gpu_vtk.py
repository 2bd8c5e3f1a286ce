"""Node positions computed on the device (ftk2h_result_copy_positions) equal the
reference VTK writer's formula (utils/vtk.hpp:26-59) as the oracle restates it
(oracle.positions), bit for bit, as does the host writer's ftk2_amd.vtk.positions, and
a file written from them reads back with those points."""
import sys

import numpy as np
import pytest

import oracle
from ftk2_amd import engine, synth, vtk

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cases():
    cp3 = synth.cp3d_config(n=18, nt=3, ngauss=6, seed=7, noise=0.02, zero_frac=0.02)
    f = synth.scalar_config(n=16, nt=3, ngauss=6, seed=8, dtype=np.float32)
    g = synth.scalar_config(n=16, nt=3, ngauss=6, seed=9, dtype=np.float32)
    u, v = synth.cp2d_config(nx=40, ny=36, nt=5, ngauss=8, seed=10)
    return [("cp3d_f64", cp3, 0, 3, (0.0, 0.0)), ("contour_f32", [f], 1, 1, (0.02, 0.0)),
            ("fiber_f32", [f, g], 2, 2, (0.0, 0.01)), ("cp2d_f64", [u, v], 0, 2, (0.0, 0.0))]


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_device_positions(case, tmp_path):
    name, comps, kind, m, thr = case
    say(f"[vtk] {name} {comps[0].shape}")
    mesh = engine.RegularSimplicialMesh(list(comps[0].shape))
    res = engine.extract(mesh, kind, m, comps, thresholds=thr)
    cx = res.to_complex()
    assert len(cx.vertices) > 0
    p = res.positions()
    want, _ = oracle.positions(cx.vertices, cx.bary, mesh.global_dims)
    assert np.array_equal(p, want)
    assert np.array_equal(p, vtk.positions(cx.vertices, cx.bary, mesh.global_dims))
    vtk.write_vtu(cx, mesh.global_dims, tmp_path / "r.vtu", points=p)
    r = vtk.read(tmp_path / "r.vtu")
    assert np.array_equal(r["Points/Points"], p)
    assert np.array_equal(r["PointData/Scalar"], cx.scalar)
