"""Streaming extraction on the GPU: appending timesteps one at a time gives the same
complex as the one-shot extraction (test_examples.cpp:259-297 contract) and as the
oracle's execute_stream restatement."""
import numpy as np
import pytest

import oracle
from ftk2_amd import _lib, engine, synth

pytestmark = pytest.mark.gpu


def run_stream(comps, kind, m, thr=(0.0, 0.0), window=2, retire=None):
    nt = comps[0].shape[-1]
    dt = _lib.F64 if comps[0].dtype == np.float64 else _lib.F32
    s = engine.Stream(comps[0].shape[:-1], kind, m, dt, len(comps), thr, window=window, retire_bytes=retire)
    for t in range(nt):
        s.append([np.asfortranarray(c[..., t]) for c in comps])
    r = s.finish().to_complex()
    s.close()
    return r


def same(a, b):
    assert np.array_equal(a.vertices, b.vertices)
    assert np.array_equal(a.lam.view(np.uint64), b.lam.view(np.uint64))
    assert np.array_equal(a.track_id, b.track_id)
    assert sorted(a.connectivity) == sorted(b.connectivity)
    for d in a.connectivity:
        assert np.array_equal(a.connectivity[d], b.connectivity[d])


# One GPU test run ended with a host segmentation fault inside the oracle's
# execute_stream restatement during the fiber case, after the GPU stream had matched
# the one-shot extraction.  It did not recur on the CPU (torch imported first, so the
# oracle runs on torch's bundled libgomp; 1, 16, 64 threads; AddressSanitizer), and
# its cause is not known, so the oracle's stream restatement is not run on the GPU
# host for fiber: the fiber stream is compared with the oracle's one-shot
# restatement instead, and tests/test_oracle_engine.py checks on the CPU that the
# stream restatement equals the one-shot one for every predicate.
ORACLE_STREAM_ON_GPU_HOST = {"cp3d", "cp2d", "contour", "cp3d_float"}


@pytest.mark.parametrize("case", ["cp3d", "cp2d", "contour", "fiber", "cp3d_float"])
def test_stream_equals_oneshot(case):
    print(f"[stream] {case}", flush=True)
    if case == "cp3d":
        comps, kind, m, thr = synth.cp3d_config(n=18, nt=6, seed=3), 0, 3, (0, 0)
    elif case == "cp3d_float":
        comps = synth.gaussians((14, 13, 12), 5, 6, sigma=(8, 24), seed=4, noise=0.02, zero_frac=0.02,
                                dup_frac=0.02, dtype=np.float32)
        kind, m, thr = 0, 3, (0, 0)
    elif case == "cp2d":
        comps, kind, m, thr = list(synth.cp2d_config(nx=40, ny=36, nt=7, seed=2)), 0, 2, (0, 0)
    elif case == "contour":
        comps = [synth.scalar_config(n=16, nt=6, ngauss=8, seed=1, dtype=np.float64)]
        kind, m, thr = 1, 1, (0.0, 0)
    else:
        comps = [synth.scalar_config(n=14, nt=5, ngauss=8, seed=s, dtype=np.float64) for s in (7, 8)]
        kind, m, thr = 2, 2, (0.0, 0.0)
    mesh = engine.RegularSimplicialMesh(comps[0].shape)
    one = engine.extract(mesh, kind, m, comps, thr).to_complex()
    st = run_stream(comps, kind, m, thr)
    same(st, one)
    # and the oracle's execute_stream (fiber: its one-shot execute, see above)
    if case in ORACLE_STREAM_ON_GPU_HOST:
        steps = [[np.asfortranarray(c[..., t]) for c in comps] for t in range(comps[0].shape[-1])]
        ref = oracle.run_stream(steps, kind, m, thr=thr)
    else:
        ref = oracle.run(comps, kind, m, thr=thr)
    assert len(ref["verts"]) > 0
    assert np.array_equal(st.vertices, ref["verts"])
    assert np.array_equal(st.track_id, ref["track"])
    for d, cells in ref["conn"].items():
        assert np.array_equal(st.connectivity[d], cells)


def test_stream_two_timesteps_and_empty():
    comps = synth.cp3d_config(n=10, nt=2, seed=5)
    mesh = engine.RegularSimplicialMesh(comps[0].shape)
    same(run_stream(comps, 0, 3), engine.extract(mesh, 0, 3, comps).to_complex())
    s = engine.Stream([10, 10, 10], 0, 3, _lib.F64, 3)
    s.append([np.asfortranarray(c[..., 0]) for c in comps])
    r = s.finish()
    assert r.sizes()[0] == 0
    s.close()


def stream_stats(comps, kind, m, thr, retire):
    nt = comps[0].shape[-1]
    dt = _lib.F64 if comps[0].dtype == np.float64 else _lib.F32
    s = engine.Stream(comps[0].shape[:-1], kind, m, dt, len(comps), thr, retire_bytes=retire)
    resident = []
    for t in range(nt):
        s.append([np.asfortranarray(c[..., t]) for c in comps])
        resident.append(s.resident_bytes())
    retired = s.retired_nodes()
    r = s.finish().to_complex()
    s.close()
    return r, resident, retired


@pytest.mark.parametrize("retire", [0, 4096])
@pytest.mark.parametrize("case", ["cp2d", "cp3d", "contour", "fiber"])
def test_stream_retirement_equals_oneshot(case, retire):
    """Finished nodes, lambda and cells moved to host memory (after every slab, or
    past a small threshold) give the one-shot complex bit for bit, tracks included
    (long-lived tracks that merge after their first nodes retired); the device then
    holds only the timestep buffers and the frontier."""
    print(f"[stream-retire] {case} {retire}", flush=True)
    if case == "cp2d":
        comps, kind, m, thr = list(synth.cp2d_config(nx=40, ny=36, nt=24, seed=12)), 0, 2, (0, 0)
    elif case == "cp3d":
        comps, kind, m, thr = synth.cp3d_config(n=16, nt=8, ngauss=6, seed=13, noise=0.02), 0, 3, (0, 0)
    elif case == "contour":
        comps = [synth.scalar_config(n=16, nt=10, ngauss=8, seed=14, dtype=np.float64)]
        kind, m, thr = 1, 1, (0.0, 0)
    else:
        comps = [synth.scalar_config(n=14, nt=8, ngauss=8, seed=s, dtype=np.float64) for s in (15, 16)]
        kind, m, thr = 2, 2, (0.0, 0.0)
    mesh = engine.RegularSimplicialMesh(comps[0].shape)
    one = engine.extract(mesh, kind, m, comps, thr).to_complex()
    assert len(one.vertices) > 50
    st, resident, retired = stream_stats(comps, kind, m, thr, retire)
    same(st, one)
    assert retired > 0
    # some track spans more than half of the timesteps, so it stayed open across retirements
    nsp = int(np.prod(comps[0].shape[:-1]))
    tmin = one.vertices.min(axis=1).astype(np.int64) // nsp
    span = max(np.ptp(tmin[one.track_id == k]) for k in np.unique(one.track_id))
    assert span > comps[0].shape[-1] // 2
    keep, res_keep, ret_keep = stream_stats(comps, kind, m, thr, 2**64 - 1)
    same(keep, one)
    assert ret_keep == 0
    assert resident[-1] < res_keep[-1]
    if retire == 0:
        assert max(resident[2:]) < res_keep[-1]


def window_case(case):
    if case == "cp2d":
        return list(synth.cp2d_config(nx=40, ny=36, nt=9, seed=21)), 0, 2, (0, 0)
    if case == "cp3d":
        return synth.cp3d_config(n=14, nt=7, ngauss=6, seed=22, noise=0.02), 0, 3, (0, 0)
    if case == "contour":
        return [synth.scalar_config(n=16, nt=8, ngauss=8, seed=23, dtype=np.float64)], 1, 1, (0.0, 0)
    comps = [synth.scalar_config(n=14, nt=7, ngauss=8, seed=s, dtype=np.float64) for s in (24, 25)]
    return comps, 2, 2, (0.0, 0.0)


@pytest.mark.parametrize("case", ["cp2d", "cp3d", "contour", "fiber"])
def test_stream_windows_equal_oneshot(case):
    """Windows of several slabs (ftk2h_stream_set_window): full windows, a last
    partial window, one window over the whole stream, and a window larger than the
    stream all give the one-shot complex bit for bit, as the slab-by-slab stream does;
    with retirement after every window too."""
    print(f"[stream-window] {case}", flush=True)
    comps, kind, m, thr = window_case(case)
    nt = comps[0].shape[-1]
    mesh = engine.RegularSimplicialMesh(comps[0].shape)
    one = engine.extract(mesh, kind, m, comps, thr).to_complex()
    assert len(one.vertices) > 50
    for w in (3, 4, nt, nt + 3):
        print(f"[stream-window] {case} window {w}", flush=True)
        same(run_stream(comps, kind, m, thr, window=w), one)
    same(run_stream(comps, kind, m, thr, window=3, retire=0), one)


def test_stream_window_is_set_before_the_first_timestep():
    comps = list(synth.cp2d_config(nx=12, ny=10, nt=3, seed=2))
    s = engine.Stream(comps[0].shape[:-1], 0, 2, _lib.F64, 2)
    s.append([np.asfortranarray(c[..., 0]) for c in comps])
    with pytest.raises(_lib.Ftk2hError) as e:
        _lib.check(_lib.lib().ftk2h_stream_set_window(s.h, 3))
    assert e.value.code == 1
    with pytest.raises(_lib.Ftk2hError):
        engine.Stream(comps[0].shape[:-1], 0, 2, _lib.F64, 2, window=1)
    s.close()


def test_stream_index_overflow_is_reported(monkeypatch):
    """Committed node indices are uint32: a stream that would pass the limit fails with
    FTK2H_E_OVERFLOW instead of wrapping (the limit lowered through
    FTK2H_STREAM_NODE_LIMIT, read when the stream opens)."""
    comps = list(synth.cp2d_config(nx=40, ny=36, nt=8, seed=2))
    mesh = engine.RegularSimplicialMesh(comps[0].shape)
    n = len(engine.extract(mesh, 0, 2, comps).to_complex().vertices)
    assert n > 40
    monkeypatch.setenv("FTK2H_STREAM_NODE_LIMIT", str(n // 2))
    s = engine.Stream(comps[0].shape[:-1], 0, 2, _lib.F64, 2)
    with pytest.raises(_lib.Ftk2hError) as e:
        for t in range(8):
            s.append([np.asfortranarray(c[..., t]) for c in comps])
        s.finish()
    assert e.value.code == 3
    s.close()
    monkeypatch.delenv("FTK2H_STREAM_NODE_LIMIT")
    same(run_stream(comps, 0, 2), engine.extract(mesh, 0, 2, comps).to_complex())
