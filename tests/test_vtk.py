"""VTK output (ftk2_amd/vtk.py and include/ftk2/utils/vtk.hpp), the reference's
write_complex_to_vtu / write_complex_to_vtp (utils/vtk.hpp:26-160): positions,
point arrays and cells, in both data modes, read back by an independent parser
(base64 + XML) and checked against the complex; the C++ writer and the Python
writer produce the same arrays."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from ftk2_amd import synth, vtk
from ftk2_amd.engine import FeatureComplex

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def complex_of(comps, kind, m, thr=(0.0, 0.0)):
    ref = oracle.run(comps, kind, m, thr=thr)
    return FeatureComplex(ref["verts"], ref["lam"], ref["bary"], ref["scalar"], ref["track"].astype(np.uint64),
                          ref["conn"])


@pytest.fixture(scope="module")
def cases():
    cp3 = synth.cp3d_config(n=10, nt=3, ngauss=4, seed=3, noise=0.05, zero_frac=0.02)
    f = synth.scalar_config(n=9, nt=3, ngauss=4, seed=4, dtype=np.float32)
    u, v = synth.cp2d_config(nx=20, ny=18, nt=4, ngauss=5, seed=5)
    return {
        "cp3d": (complex_of(cp3, oracle.PRED_CP, 3), list(cp3[0].shape)),
        "contour": (complex_of([f], oracle.PRED_CONTOUR, 1, (0.01, 0.0)), list(f.shape)),
        "cp2d": (complex_of([u, v], oracle.PRED_CP, 2), list(u.shape)),
    }


def test_positions_properties(cases):
    cx, gd = cases["cp3d"]
    assert len(cx.vertices) > 50
    p = vtk.positions(cx.vertices, cx.bary, gd)
    # coordinates of every vertex (x fastest over the global dims)
    co = np.stack(np.unravel_index(cx.vertices.astype(np.int64), gd, order="F"), -1).astype(np.float64)
    lo, hi = co.min(1)[:, :3], co.max(1)[:, :3]
    # inside the simplex's bounding box, up to the solver's 1e-9 slack and float rounding
    assert np.all(p >= lo - 1e-5) and np.all(p <= hi + 1e-5)
    # a node at a vertex (one coordinate 1, the rest 0) sits exactly on that vertex
    unit = np.isclose(cx.bary, 1.0) & (cx.bary == 1.0)
    rows = np.nonzero(unit.any(1) & (np.count_nonzero(cx.bary, 1) == 1))[0]
    for r in rows:
        k = int(np.argmax(cx.bary[r]))
        assert np.array_equal(p[r], co[r, k, :3].astype(np.float32))
    # an exact centroid lands on the mean of the vertices
    b = np.full((1, 4), 0.25, np.float32)
    verts = np.array([[0, 1, 1 + 5, 1 + 5 + 25]], np.uint64)
    assert np.array_equal(vtk.positions(verts, b, [5, 5, 5, 3])[0], np.array([0.75, 0.5, 0.25], np.float32))


def check_file(path, cx, gd, kind, target_dim=-1):
    r = vtk.read(path)
    n = len(cx.vertices)
    assert int(r["piece"]["NumberOfPoints"]) == n
    assert np.array_equal(r["Points/Points"].reshape(-1, 3), vtk.positions(cx.vertices, cx.bary, gd))
    assert np.array_equal(r["PointData/Scalar"], cx.scalar.astype(np.float32))
    assert np.array_equal(r["PointData/TrackID"], cx.track_id.astype(np.int64).astype(np.int32))
    assert np.array_equal(r["PointData/Type"], np.zeros(n, np.int32))
    assert np.array_equal(r["PointData/Time"], np.zeros(n, np.float32))
    if kind == "vtu":
        conn, offs, types = r["Cells/connectivity"], r["Cells/offsets"], r["Cells/types"]
        if target_dim == 0:
            assert np.array_equal(conn, np.arange(n)) and np.all(types == 1)
            return
        dims = [d for d in sorted(cx.connectivity) if target_dim in (-1, d) and d <= 3]
        exp = np.concatenate([cx.connectivity[d].ravel() for d in dims]) if dims else np.zeros(0)
        assert np.array_equal(conn, exp.astype(np.int64))
        sizes = np.concatenate([np.full(len(cx.connectivity[d]), d + 1) for d in dims]) if dims else np.zeros(0)
        assert np.array_equal(offs, np.cumsum(sizes))
        assert np.array_equal(types, np.array([vtk.VTK_TYPE[int(s) - 1] for s in sizes], np.uint8))
        assert int(r["piece"]["NumberOfCells"]) == len(types)
    else:
        for d, sec in ((0, "Verts"), (1, "Lines"), (2, "Polys")):
            c = cx.connectivity.get(d) if target_dim in (-1, d) else None
            c = np.zeros((0, d + 1), np.int64) if c is None else c
            assert np.array_equal(r[f"{sec}/connectivity"], c.ravel().astype(np.int64))
            assert np.array_equal(r[f"{sec}/offsets"], (d + 1) * np.arange(1, len(c) + 1))


@pytest.mark.parametrize("name", ["cp3d", "contour", "cp2d"])
@pytest.mark.parametrize("ascii_mode", [False, True])
def test_python_writer_roundtrip(cases, tmp_path, name, ascii_mode):
    cx, gd = cases[name]
    assert len(cx.vertices) > 0
    vtk.write_vtu(cx, gd, tmp_path / "a.vtu", ascii=ascii_mode)
    check_file(tmp_path / "a.vtu", cx, gd, "vtu")
    vtk.write_vtp(cx, gd, tmp_path / "a.vtp", ascii=ascii_mode)
    check_file(tmp_path / "a.vtp", cx, gd, "vtp")
    for td in (0, 1, 3):
        vtk.write_vtu(cx, gd, tmp_path / f"t{td}.vtu", target_dim=td, ascii=ascii_mode)
        check_file(tmp_path / f"t{td}.vtu", cx, gd, "vtu", td)


def dump(cx, gd, path):
    m = cx.vertices.shape[1] - 1
    with open(path, "w") as f:
        f.write(f"{len(gd)} {' '.join(map(str, gd))} {m} {len(cx.vertices)}\n")
        for i in range(len(cx.vertices)):
            f.write(" ".join(str(int(x)) for x in cx.vertices[i]) + " ")
            f.write(" ".join(repr(float(x)) for x in cx.bary[i]) + " ")
            f.write(f"{float(cx.scalar[i])!r} {int(cx.track_id[i])}\n")
        f.write(f"{len(cx.connectivity)}\n")
        for d in sorted(cx.connectivity):
            c = cx.connectivity[d]
            f.write(f"{d} {len(c)} " + " ".join(str(int(x)) for x in c.ravel()) + "\n")


@pytest.mark.parametrize("name", ["cp3d", "contour", "cp2d"])
def test_cpp_writer_matches(cases, tmp_path, name):
    cx, gd = cases[name]
    exe = tmp_path / "vtk_write"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "vtk_write.cpp"), "-o", str(exe)], check=True)
    dump(cx, gd, tmp_path / "cx.txt")
    subprocess.run([str(exe), str(tmp_path / "cx.txt"), str(tmp_path / "c")], check=True, capture_output=True)
    for mode in ("ascii", "bin"):
        check_file(tmp_path / f"c_{mode}.vtu", cx, gd, "vtu")
        check_file(tmp_path / f"c_{mode}.vtp", cx, gd, "vtp")
    check_file(tmp_path / "c_dim0.vtu", cx, gd, "vtu", 0)
    check_file(tmp_path / "c_dim1.vtp", cx, gd, "vtp", 1)
