"""The oracle on an ExtrudedSimplicialMesh over a regular base (hguo/ftk2
include/ftk2/core/mesh.hpp:542-717).  Its simplices and d-cells are the Kuhn
simplices of the regular mesh with one more axis (a prism j-subdivision of a base
(k-1)-simplex x [t, t+1] is the chain that gains the time step after its j-th
vertex), and get_value reads the same array element (coordinates = base
coordinates + layer); only the vertex ids differ: id = spatial id + t * N with
N = n_spatial_verts_ = grid_index_to_id(base global dims) (mesh.hpp:547-552),
larger than the number of spatial vertices.

Pins:
* N for a 2-D and a 3-D base, from the reference's formula evaluated by hand;
* critical points and fibers: the id change preserves order, and SoS uses ids
  only through their order (det2 ties; det3 ignores them), so nodes, lambda,
  cells and tracks equal the regular mesh's under the id map;
* contour: sign ties (a value exactly at the threshold) take the parity of the
  extruded id (sos.hpp:12-20) - checked by brute force over every edge of a
  small grid with planted ties, in a case where the two id schemes disagree.
"""
import numpy as np

import oracle
from ftk2_amd import synth


def ext_to_reg(ids, spatial, N):
    ids = np.asarray(ids, dtype=np.uint64)
    t = ids // np.uint64(N)
    sp = ids % np.uint64(N)
    return sp + t * np.uint64(int(np.prod(spatial)))


def test_time_stride_formula():
    # grid_index_to_id(g) over global dims g: g0 + g1*g0 + g2*g0*g1
    assert oracle.extruded_time_stride([7, 5]) == 7 + 5 * 7
    assert oracle.extruded_time_stride([4, 3, 6]) == 4 + 3 * 4 + 6 * 12
    assert oracle.extruded_time_stride([7, 5]) > 7 * 5


def same_under_map(reg, ext, spatial, N):
    assert np.array_equal(ext_to_reg(ext["verts"], spatial, N), reg["verts"].astype(np.uint64))
    assert np.array_equal(ext["lam"], reg["lam"])
    assert np.array_equal(ext["track"], reg["track"])
    assert sorted(ext["conn"]) == sorted(reg["conn"])
    for d in reg["conn"]:
        assert np.array_equal(ext["conn"][d], reg["conn"][d])


def test_cp2d_and_cp3d_equal_regular_under_id_map():
    u, v = synth.gaussians((13, 11), 5, 5, seed=21, zero_frac=0.03, dup_frac=0.03)
    N = oracle.extruded_time_stride([13, 11])
    reg = oracle.run([u, v], oracle.PRED_CP, 2)
    ext = oracle.run([u, v], oracle.PRED_CP, 2, time_stride=N)
    assert len(reg["verts"]) > 20
    same_under_map(reg, ext, [13, 11], N)
    c = synth.cp3d_config(n=8, nt=3, ngauss=3, seed=22, noise=0.03, zero_frac=0.03)
    N3 = oracle.extruded_time_stride([8, 8, 8])
    reg = oracle.run(c, oracle.PRED_CP, 3)
    ext = oracle.run(c, oracle.PRED_CP, 3, time_stride=N3)
    assert len(reg["verts"]) > 20
    same_under_map(reg, ext, [8, 8, 8], N3)


def test_fiber_equal_regular_under_id_map():
    a = synth.gaussians((12, 10, 9), 3, 4, gradient=False, seed=23, subtract_mean=True)[0]
    b = synth.gaussians((12, 10, 9), 3, 4, gradient=False, seed=24, subtract_mean=True)[0]
    N = oracle.extruded_time_stride([12, 10, 9])
    reg = oracle.run([a, b], oracle.PRED_FIBER, 2)
    ext = oracle.run([a, b], oracle.PRED_FIBER, 2, time_stride=N)
    assert len(reg["verts"]) > 20
    same_under_map(reg, ext, [12, 10, 9], N)


def test_contour_ties_follow_extruded_parity():
    nx, ny, nt = 7, 4, 3  # N = 7 + 4*7 = 35 (odd), nx*ny = 28 (even): parities differ at odd t
    rng = np.random.default_rng(5)
    f = np.asfortranarray(rng.uniform(-1, 1, size=(nx, ny, nt)))
    ties = [(1, 1, 1), (3, 2, 1), (5, 0, 1), (2, 3, 2), (4, 1, 0)]
    for p in ties:
        f[p] = 0.0
    N = oracle.extruded_time_stride([nx, ny])
    assert N % 2 != (nx * ny) % 2
    ext = oracle.run([f], oracle.PRED_CONTOUR, 1, time_stride=N)
    reg = oracle.run([f], oracle.PRED_CONTOUR, 1)

    def brute(idf):
        q = np.trunc(f * 1e6 + np.where(f > 0, 0.5, -0.5))
        nodes = set()
        for e in oracle.chains(3, 1):
            dx = [(int(e[1]) >> j) & 1 for j in range(3)]
            for t in range(nt - dx[2]):
                for y in range(ny - dx[1]):
                    for x in range(nx - dx[0]):
                        a, b = (x, y, t), (x + dx[0], y + dx[1], t + dx[2])
                        ia, ib = idf(a), idf(b)
                        sa = 1 if q[a] >= 1 else (-1 if q[a] <= -1 else (1 if ia % 2 == 0 else -1))
                        sb = 1 if q[b] >= 1 else (-1 if q[b] <= -1 else (1 if ib % 2 == 0 else -1))
                        if sa != sb:
                            nodes.add((min(ia, ib), max(ia, ib)))
        return nodes

    ext_ids = lambda g: g[0] + nx * g[1] + N * g[2]
    reg_ids = lambda g: g[0] + nx * g[1] + nx * ny * g[2]
    assert {tuple(int(x) for x in r) for r in ext["verts"]} == brute(ext_ids)
    assert {tuple(int(x) for x in r) for r in reg["verts"]} == brute(reg_ids)
    # the planted ties at odd t flip sign between the schemes, so the node sets differ
    mapped = {tuple(int(x) for x in r) for r in ext_to_reg(ext["verts"], [nx, ny], N)}
    assert mapped != {tuple(int(x) for x in r) for r in reg["verts"]}
