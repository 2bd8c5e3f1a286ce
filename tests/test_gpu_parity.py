"""GPU parity: the HIP library (through the C ABI) against the CPU oracle on the same
seeded inputs.  Bar: node simplices, float barycentrics, track ids and connectivity
bit-exact; FP64 lambda within 1e-12 relative (FP32 within 1e-5)."""
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


def compare(name, comps, kind, m, thr=(0.0, 0.0), offset=None, gdims=None, device_inputs=False):
    say(f"[parity] {name}: shape={comps[0].shape} dtype={comps[0].dtype} kind={kind} m={m}")
    ref = oracle.run(comps, KIND[kind], m, thr=thr, offset=offset, gdims=gdims)
    mesh = engine.RegularSimplicialMesh(comps[0].shape, offset, gdims)
    inputs = comps
    if device_inputs:
        import torch
        inputs = [torch.from_numpy(np.ascontiguousarray(c.transpose())).cuda().t() if c.ndim == 2
                  else torch.from_numpy(np.ascontiguousarray(c.transpose())).cuda().permute(
                      *reversed(range(c.ndim))) for c in comps]
    res = engine.extract(mesh, KIND[kind], m, inputs, thr)
    cx = res.to_complex()
    n = len(ref["verts"])
    say(f"[parity] {name}: oracle nodes={n} gpu nodes={len(cx.vertices)}")
    assert np.array_equal(cx.vertices, ref["verts"]), "node simplices differ"
    tol = 1e-12 if comps[0].dtype == np.float64 else 1e-5
    if n:
        den = np.maximum(np.abs(ref["lam"]), 1e-300)
        rel = np.max(np.abs(cx.lam - ref["lam"]) / den)
        assert rel <= tol, f"lambda rel diff {rel}"
    assert np.array_equal(cx.bary, ref["bary"]), "float barycentrics differ"
    assert np.array_equal(cx.scalar, ref["scalar"]), "FeatureElement::scalar differs"
    assert np.array_equal(cx.track_id, ref["track"]), "track ids differ"
    assert sorted(cx.connectivity) == sorted(ref["conn"])
    for dim, cells in ref["conn"].items():
        assert np.array_equal(cx.connectivity[dim], cells), f"connectivity dim {dim} differs"
    return cx, ref


def test_cp3d_linear_float_reference_case():
    compare("cp3d_linear_float_24^3x8", refcases.cp3d_linear_float(), "cp", 3)


def test_levelset3d_reference_case():
    cx, _ = compare("levelset3d_float_32^3x10", refcases.levelset3d_float(), "contour", 1)
    assert len(cx.vertices) == 66418


def test_cp2d_woven_reference_case():
    """test_examples.cpp:39-105: offset mesh, separate U/V arrays, scalar variable."""
    u, v, s = refcases.woven_cp2d()
    say("[parity] cp2d_woven_16^2x5 offset (1,1,0)")
    mesh = engine.RegularSimplicialMesh([14, 14, 5], [1, 1, 0], [16, 16, 5])
    pred = engine.CriticalPointPredicate(M=2, var_names=["U", "V"], scalar_var_name="Woven")
    eng = engine.SimplicialEngine(mesh, pred)
    eng.execute({"U": u, "V": v, "Woven": s})
    cx = eng.get_complex()
    ref = oracle.run([u, v], oracle.PRED_CP, 2, dims=(14, 14, 5), offset=(1, 1, 0), gdims=(16, 16, 5), scalar=s)
    assert len(cx.vertices) == 788
    assert np.array_equal(cx.vertices, ref["verts"])
    assert np.array_equal(cx.lam.view(np.uint64), ref["lam"].view(np.uint64))
    assert np.array_equal(cx.scalar, ref["scalar"])
    assert np.array_equal(cx.track_id, ref["track"])
    assert sorted(cx.connectivity) == sorted(ref["conn"])
    for dim, cells in ref["conn"].items():
        assert np.array_equal(cx.connectivity[dim], cells)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cp2d_config0(seed):
    u, v = synth.cp2d_config(seed=seed)
    compare(f"cp2d_64x64x8_seed{seed}", [u, v], "cp", 2)


@pytest.mark.parametrize("seed,n", [(0, 24), (1, 32), (2, 20)])
def test_cp3d_degenerate_double(seed, n):
    comps = synth.cp3d_config(n=n, nt=4, seed=seed)
    compare(f"cp3d_{n}^3x4_seed{seed}", comps, "cp", 3)


def test_cp3d_noise_heavy_float():
    comps = synth.gaussians((20, 18, 16), 4, 6, sigma=(8, 24), seed=7, noise=1e-3, zero_frac=0.02,
                            dup_frac=0.05, dtype=np.float32)
    compare("cp3d_float_noise_20x18x16x4", comps, "cp", 3)


@pytest.mark.parametrize("sigma,dtype", [((1.5, 2.5), np.float64), ((2.0, 3.0), np.float64),
                                         ((1.5, 2.5), np.float32)])
def test_cp3d_zero_regions(sigma, dtype):
    """Narrow Gaussians: large regions whose values round to zero (12-28% of the
    vertices), some below the solver's pivot threshold.  Their tets are inside by SoS
    and the solver gate decides them."""
    comps = synth.gaussians((24, 22, 20), 4, 4, sigma=sigma, seed=5, dtype=dtype)
    compare(f"cp3d zero regions {sigma} {np.dtype(dtype).name}", comps, "cp", 3)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cp3d_sub_quantum_linear_field(dtype):
    """v = 1e-9 A (x - c(t)): every value rounds to zero (|v| Q < 1/2), so every tet
    is inside by SoS, and the solver gate (pivots ~1e-9 above its 1e-12 threshold)
    keeps exactly the tets around the moving zero: the nodes all lie in the all-zero
    region."""
    A = [[1.0, 0.3, -0.2], [0.1, -1.5, 0.4], [0.2, 0.1, 2.0]]
    dims, nt, c0, vel = (12, 11, 10), 4, (5.37, 4.21, 4.83), (0.27, -0.19, 0.11)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in dims], np.arange(nt, dtype=np.float64),
                        indexing="ij")
    rel = [grids[j] - (c0[j] + vel[j] * grids[3]) for j in range(3)]
    comps = [np.asfortranarray((1e-9 * sum(A[i][j] * rel[j] for j in range(3))).astype(dtype)) for i in range(3)]
    assert max(np.abs(c).max() for c in comps) * 1e6 < 0.5
    cx, ref = compare(f"cp3d sub-quantum linear {np.dtype(dtype).name}", comps, "cp", 3)
    assert len(ref["verts"]) > 0


def test_cp3d_large_magnitudes_exact_path():
    # |v| > 1.05 -> quantized values beyond the fast filter: verbatim exact path
    comps = synth.gaussians((14, 12, 10), 3, 5, sigma=(3, 6), amp=(20.0, 40.0), seed=9)
    compare("cp3d_large_magnitude", comps, "cp", 3)


def test_contour4d_double_threshold():
    s = synth.scalar_config(n=24, nt=4, ngauss=12, seed=3, dtype=np.float64)
    compare("contour_24^3x4_thr", [s], "contour", 1, thr=(0.01, 0.0))


def test_contour3d_marching_tet():
    s = synth.gaussians((30, 28), 6, 8, gradient=False, seed=4, subtract_mean=True)[0]
    compare("contour_2d+t", [s], "contour", 1)


def test_offset_slab_mesh():
    comps = synth.cp3d_config(n=16, nt=6, seed=4)
    slab = [np.asfortranarray(c[:, :, :, 2:4]) for c in comps]
    compare("cp3d_offset_slab", slab, "cp", 3, offset=[0, 0, 0, 2], gdims=[16, 16, 16, 6])


def test_device_inputs_in_place():
    comps = synth.cp3d_config(n=18, nt=3, seed=6)
    compare("cp3d_torch_device_inputs", comps, "cp", 3, device_inputs=True)


def test_cp1d_engine_case():
    compare("cp1d_engine", refcases.cp1d_engine(), "cp", 1)


def test_empty_mesh_dimension_one():
    u = np.zeros((5, 1, 4), order="F")
    mesh = engine.RegularSimplicialMesh(u.shape)
    res = engine.extract(mesh, oracle.PRED_CP, 2, [u, u])
    assert res.sizes()[0] == 0


def test_overflow_capacity_reported():
    comps = synth.cp3d_config(n=16, nt=3, seed=1)
    mesh = engine.RegularSimplicialMesh(comps[0].shape)
    with pytest.raises(_lib.Ftk2hError) as e:
        engine.extract(mesh, oracle.PRED_CP, 3, comps, node_capacity=1)
    assert e.value.code == 3


def test_run_twice_bit_identical():
    comps = synth.cp3d_config(n=28, nt=4, seed=8)
    mesh = engine.RegularSimplicialMesh(comps[0].shape)
    a = engine.extract(mesh, oracle.PRED_CP, 3, comps).to_complex()
    b = engine.extract(mesh, oracle.PRED_CP, 3, comps).to_complex()
    assert np.array_equal(a.vertices, b.vertices)
    assert np.array_equal(a.lam.view(np.uint64), b.lam.view(np.uint64))
    assert np.array_equal(a.track_id, b.track_id)
    assert np.array_equal(a.connectivity[1], b.connectivity[1])


def test_emit_overflow_paths(monkeypatch):
    """Workgroups whose nodes or cells exceed the LDS staging capacity take the
    per-thread path; results must not change."""
    comps = synth.cp3d_config(n=22, nt=4, seed=12)
    monkeypatch.setenv("FTK2H_EMIT_CAPS", "16,8")
    compare("cp3d_emit_overflow", comps, "cp", 3)
    s = synth.scalar_config(n=20, nt=4, ngauss=10, seed=2, dtype=np.float32)
    compare("contour_emit_overflow", [s], "contour", 1)
    a = synth.scalar_config(n=20, nt=4, ngauss=10, seed=5, dtype=np.float32)
    b = synth.scalar_config(n=20, nt=4, ngauss=10, seed=6, dtype=np.float32)
    compare("fiber_emit_overflow", [a, b], "fiber", 2)
    t = synth.gaussians((26, 24), 5, 6, gradient=False, seed=8, subtract_mean=True)[0]
    compare("contour_2d+t_emit_overflow", [t], "contour", 1)


def test_fiber3d_reference_case():
    cx, _ = compare("fiber3d_double_32^3x10", refcases.fiber3d_double(), "fiber", 2)
    assert len(cx.vertices) == 6078


def test_fiber_float_config_like():
    a = synth.scalar_config(n=20, nt=4, ngauss=10, seed=5, dtype=np.float32)
    b = synth.scalar_config(n=20, nt=4, ngauss=10, seed=6, dtype=np.float32)
    compare("fiber_float_20^3x4", [a, b], "fiber", 2)


def test_fiber_on_3d_mesh_star_edges():
    a = synth.gaussians((26, 24), 5, 6, gradient=False, seed=8, subtract_mean=True)[0]
    b = synth.gaussians((26, 24), 5, 6, gradient=False, seed=9, subtract_mean=True)[0]
    compare("fiber_2d+t", [a, b], "fiber", 2, thr=(0.001, -0.002))


def test_cp2_on_4d_mesh_star_fan():
    u, v = synth.gaussians((12, 11, 10), 3, 5, sigma=(3, 6), seed=10)[:2]
    compare("cp2_on_3d+t", [u, v], "cp", 2)
