"""Batched extraction (ftk2h_batch_*): every job's result equals ftk2h_extract_regular
on that job alone, bit for bit (node simplices, lambda, barycentrics, track ids,
connectivity), for every schedule, with and without graph replay; a few jobs are
also checked against the oracle."""
import sys

import numpy as np
import pytest

import oracle
from ftk2_amd import engine, synth
from ftk2_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b, what):
    assert np.array_equal(a.vertices, b.vertices), f"{what}: node simplices differ"
    assert np.array_equal(a.lam, b.lam), f"{what}: lambda differs"
    assert np.array_equal(a.bary, b.bary), f"{what}: barycentrics differ"
    assert np.array_equal(a.scalar, b.scalar), f"{what}: scalar differs"
    assert np.array_equal(a.track_id, b.track_id), f"{what}: track ids differ"
    assert sorted(a.connectivity) == sorted(b.connectivity), f"{what}: connectivity dims differ"
    for d in a.connectivity:
        assert np.array_equal(a.connectivity[d], b.connectivity[d]), f"{what}: connectivity differs"


def cp2d_jobs(n, seed=0, device=False):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        nx = int(rng.integers(2, 70))
        ny = int(rng.integers(2, 70))
        nt = int(rng.integers(2, 6))
        if i == 3:
            nx = 1  # no hypercube: empty result
        comps = synth.gaussians((nx, ny), nt, int(rng.integers(4, 16)), sigma=(4.0, 10.0), seed=seed * 1000 + i)
        if device:
            import torch
            comps = [torch.from_numpy(np.ascontiguousarray(c.transpose())).cuda().permute(2, 1, 0) for c in comps]
        jobs.append((engine.RegularSimplicialMesh([nx, ny, nt]), comps))
    return jobs


def plain(jobs, kind, m, thr=(0.0, 0.0)):
    return [engine.extract(j[0], kind, m, j[1], thr).to_complex() for j in jobs]


def check_batch(batch, jobs, ref, what):
    res = batch.run(jobs)
    assert len(res) == len(jobs)
    for i, (r, e) in enumerate(zip(res, ref)):
        same(r.to_complex(), e, f"{what} job {i}")
    return res


@pytest.mark.parametrize("cfg", [(0, 0, True), (0, 0, False), (2, 0, True), (4, 1, True), (4, 16, True),
                                 (1, 4, False)])
def test_batch_cp2d_equals_plain(cfg):
    say(f"[batch] cp2d schedule grid_per_cu={cfg[0]} chunk={cfg[1]} graph={cfg[2]}")
    jobs = cp2d_jobs(24, seed=1, device=(cfg[0] == 2))
    ref = plain(jobs, _lib.PRED_CRITICAL_POINT, 2)
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    b.set_config(*cfg)
    check_batch(b, jobs, ref, "first run")
    check_batch(b, jobs, ref, "replay")  # same shape: graphs replayed
    b.close()


def test_batch_against_oracle():
    jobs = cp2d_jobs(6, seed=2)
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    res = b.run(jobs)
    for i, (job, r) in enumerate(zip(jobs, res)):
        ref = oracle.run(job[1], oracle.PRED_CP, 2)
        cx = r.to_complex()
        say(f"[batch] oracle job {i}: shape={job[1][0].shape} nodes={len(ref['verts'])}")
        assert np.array_equal(cx.vertices, ref["verts"])
        assert np.array_equal(cx.bary, ref["bary"])
        assert np.array_equal(cx.track_id, ref["track"])
        for d, cells in ref["conn"].items():
            assert np.array_equal(cx.connectivity[d], cells)


def test_batch_results_outlive_next_run():
    jobs_a = cp2d_jobs(8, seed=3)
    jobs_b = cp2d_jobs(5, seed=4)
    ref_a = plain(jobs_a, _lib.PRED_CRITICAL_POINT, 2)
    ref_b = plain(jobs_b, _lib.PRED_CRITICAL_POINT, 2)
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    ra = b.run(jobs_a)
    rb = b.run(jobs_b)  # ra still alive: a new arena
    for i in (0, 5, 2):
        same(ra[i].to_complex(), ref_a[i], f"a{i}")
    del ra
    rc = b.run(jobs_a)
    for i in range(len(jobs_b)):
        same(rb[i].to_complex(), ref_b[i], f"b{i}")
    for i in range(len(jobs_a)):
        same(rc[i].to_complex(), ref_a[i], f"c{i}")


def test_batch_cp3d_with_degeneracies_and_offsets():
    jobs = []
    for i, (n, nt) in enumerate([(9, 3), (12, 2), (7, 4), (16, 2)]):
        comps = synth.cp3d_config(n=n, nt=nt, ngauss=4, seed=10 + i, zero_frac=0.03, dup_frac=0.03)
        if i == 1:  # a slab of a larger mesh
            mesh = engine.RegularSimplicialMesh([n, n, n, nt], [0, 0, 0, 5], [n, n, n, nt + 5])
        else:
            mesh = engine.RegularSimplicialMesh([n, n, n, nt])
        jobs.append((mesh, comps))
    ref = [engine.extract(j[0], _lib.PRED_CRITICAL_POINT, 3, j[1]).to_complex() for j in jobs]
    for cfg in [(0, 0, True), (4, 2, True)]:
        b = engine.Batch(_lib.PRED_CRITICAL_POINT, 3, _lib.F64, 4)
        b.set_config(*cfg)
        check_batch(b, jobs, ref, f"cp3d {cfg}")


def test_batch_cp2d_scalar_field():
    jobs = []
    for i in range(4):
        u, v = synth.cp2d_config(nx=20 + 7 * i, ny=18, nt=4, ngauss=6, seed=20 + i)
        s = synth.gaussians((20 + 7 * i, 18), 4, 6, seed=40 + i, gradient=False)[0]
        jobs.append((engine.RegularSimplicialMesh(list(u.shape)), [u, v], s))
    ref = [engine.extract(j[0], _lib.PRED_CRITICAL_POINT, 2, j[1], scalar=j[2]).to_complex() for j in jobs]
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    check_batch(b, jobs, ref, "scalar")


def test_batch_contour_and_fiber():
    jobs = []
    for i in range(5):
        n = 10 + 3 * i
        f = synth.gaussians((n, n + 1), 5, 8, seed=50 + i, gradient=False, subtract_mean=True,
                            dtype=np.float32)[0]
        jobs.append((engine.RegularSimplicialMesh(list(f.shape)), [f]))
    ref = plain(jobs, _lib.PRED_CONTOUR, 1, (0.01, 0.0))
    b = engine.Batch(_lib.PRED_CONTOUR, 1, _lib.F32, 3, thresholds=(0.01, 0.0))
    check_batch(b, jobs, ref, "contour d3")
    fjobs = []
    for i in range(4):
        n = 8 + 2 * i
        f = synth.gaussians((n, n + 2), 4, 8, seed=60 + i, gradient=False, subtract_mean=True)[0]
        g = synth.gaussians((n, n + 2), 4, 8, seed=70 + i, gradient=False, subtract_mean=True)[0]
        fjobs.append((engine.RegularSimplicialMesh(list(f.shape)), [f, g]))
    # fiber on a 3-D mesh (d = 3, m = 2): star edges
    ref = plain(fjobs, _lib.PRED_FIBER, 2, (0.0, 0.0))
    b = engine.Batch(_lib.PRED_FIBER, 2, _lib.F64, 3)
    check_batch(b, fjobs, ref, "fiber d3")


def test_batch_rejects_unbatched_and_bad_jobs():
    with pytest.raises(_lib.Ftk2hError) as e:
        engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 4)  # star-fan cells
    assert e.value.code == 2
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    u, v = synth.cp2d_config(nx=8, ny=8, nt=3)
    with pytest.raises(_lib.Ftk2hError):
        b.run([(engine.RegularSimplicialMesh([8, 8, 3, 1]), [u, v])])  # ndims differs
    assert b.run([]) == []


def test_batch_autotune_keeps_results():
    jobs = cp2d_jobs(16, seed=5)
    ref = plain(jobs, _lib.PRED_CRITICAL_POINT, 2)
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    cfg, ms = b.autotune(jobs, reps=1)
    say(f"[batch] autotune picked {cfg} ({ms:.3f} ms)")
    assert b.config() == cfg
    check_batch(b, jobs, ref, "tuned")
