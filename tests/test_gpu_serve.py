"""Serving mode (ftk2h_server_* + ftk2_amd/serve.py): requests of mixed sizes,
timestep counts and templates served together by continuous batching; each result
equals the same request processed alone (ftk2h_stream, one slab per append), bit for
bit, also when the device-memory budget makes requests wait for admission."""
import sys

import numpy as np
import pytest

import oracle
from ftk2_amd import engine, serve, synth
from ftk2_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b, what):
    assert np.array_equal(a.vertices, b.vertices), f"{what}: node simplices differ"
    assert np.array_equal(a.lam, b.lam), f"{what}: lambda differs"
    assert np.array_equal(a.bary, b.bary), f"{what}: barycentrics differ"
    assert np.array_equal(a.track_id, b.track_id), f"{what}: track ids differ"
    assert sorted(a.connectivity) == sorted(b.connectivity), f"{what}: connectivity dims differ"
    for d in a.connectivity:
        assert np.array_equal(a.connectivity[d], b.connectivity[d]), f"{what}: connectivity differs"


def mixed_requests(n, seed, device=False, arrival_rate=None):
    rng = np.random.default_rng(seed)
    reqs = []
    arr = serve.poisson_arrivals(n, arrival_rate, rng) if arrival_rate else np.zeros(n)
    for i in range(n):
        T = int(rng.integers(1, 7))
        if rng.random() < 0.7:
            side = int(rng.integers(8, 40))
            dt = _lib.F64 if rng.random() < 0.8 else _lib.F32
            comps = synth.gaussians((side, side + 3), T, 6, seed=seed * 100 + i,
                                    dtype=np.float64 if dt == _lib.F64 else np.float32)
            spatial, m = (side, side + 3), 2
        else:
            side = int(rng.integers(5, 10))
            comps = synth.cp3d_config(n=side, nt=T, ngauss=4, seed=seed * 100 + i, zero_frac=0.02,
                                      dup_frac=0.02)
            spatial, m, dt = (side, side, side), 3, _lib.F64
        steps = []
        for t in range(T):
            sl = [c[..., t] for c in comps]
            if device:
                import torch
                sl = [torch.from_numpy(np.ascontiguousarray(s.transpose())).cuda().permute(
                    *reversed(range(s.ndim))) for s in sl]
            steps.append(sl)
        hc = int(np.prod([s - 1 for s in spatial])) * max(0, T - 1)
        reqs.append(serve.Request(i, float(arr[i]), spatial, _lib.PRED_CRITICAL_POINT, m, dt, steps,
                                  simplices=hc * (12 if m == 2 else 60), meta={"comps": comps}))
    return reqs


def alone(rq):
    st = engine.Stream(rq.spatial, rq.kind, rq.m, rq.dtype, len(rq.timesteps[0]), thresholds=rq.thresholds)
    for comps in rq.timesteps:
        st.append(comps)
    res = st.finish().to_complex()
    st.close()
    return res


@pytest.mark.parametrize("device", [False, True])
def test_served_equals_alone(device):
    reqs = mixed_requests(14, seed=1, device=device)
    srv = engine.Server(64 << 30)
    outs, wall, rej = serve.run(srv, reqs, keep_results=True)
    assert rej == []
    assert len(outs) == len(reqs)
    for o in outs:
        rq = reqs[o.rid]
        say(f"[serve] request {o.rid}: spatial={rq.spatial} T={len(rq.timesteps)} m={rq.m}")
        same(o.result.to_complex(), alone(rq), f"request {o.rid}")
    assert srv.stats()["reserved_bytes"] == 0


def test_budget_makes_requests_wait_results_unchanged():
    reqs = mixed_requests(10, seed=2, arrival_rate=1000.0)
    # room for about two of the largest requests at a time
    need = []
    for rq in reqs:
        probe = engine.Server(64 << 30)
        probe.open(rq.spatial, rq.kind, rq.m, rq.dtype, len(rq.timesteps[0]), len(rq.timesteps))
        need.append(probe.stats()["reserved_bytes"])
        probe.close()
    srv = engine.Server(int(2.3 * max(need)))
    outs, wall, rej = serve.run(srv, reqs, keep_results=True)
    assert rej == []
    assert len(outs) == len(reqs)
    waited = sum(1 for o in outs if o.admitted > o.arrival + 1e-4)
    say(f"[serve] {waited} of {len(outs)} requests waited for admission; rejected {rej}")
    for o in outs:
        same(o.result.to_complex(), alone(reqs[o.rid]), f"request {o.rid}")


def test_served_equals_oneshot_and_oracle():
    reqs = mixed_requests(4, seed=3)
    srv = engine.Server(64 << 30)
    outs, _, _ = serve.run(srv, reqs, keep_results=True)
    for o in outs:
        rq = reqs[o.rid]
        comps = rq.meta["comps"]
        if len(rq.timesteps) < 2:
            assert len(o.result.to_complex().vertices) == 0  # execute_stream: fewer than 2 timesteps
            continue
        cx = o.result.to_complex()
        mesh = engine.RegularSimplicialMesh(list(comps[0].shape))
        one = engine.extract(mesh, rq.kind, rq.m, comps).to_complex()
        same(cx, one, f"request {o.rid} vs one-shot")
        ref = oracle.run(comps, oracle.PRED_CP, rq.m)
        assert np.array_equal(cx.vertices, ref["verts"])
        assert np.array_equal(cx.track_id, ref["track"])


def test_server_rejects_and_reports():
    srv = engine.Server(1 << 20)
    with pytest.raises(_lib.Ftk2hError) as e:
        srv.open((512, 512), _lib.PRED_CRITICAL_POINT, 2, _lib.F64, 2, 8)  # never fits
    assert e.value.code == 3
    sid = srv.open((8, 8), _lib.PRED_CRITICAL_POINT, 2, _lib.F64, 2, 3)
    with pytest.raises(_lib.Ftk2hError) as e:
        srv.take(sid)  # not finished
    assert e.value.code == _lib.E_AGAIN
    u, v = synth.cp2d_config(nx=8, ny=8, nt=3)
    for t in range(3):
        srv.append(sid, [u[:, :, t], v[:, :, t]])
    with pytest.raises(_lib.Ftk2hError):
        srv.append(sid, [u[:, :, 0], v[:, :, 0]])  # beyond max_timesteps
    srv.close_stream(sid)
    done = []
    while not done:
        done = srv.step()
    assert done == [sid]
    r = srv.take(sid).to_complex()
    ref = oracle.run([u, v], oracle.PRED_CP, 2)
    assert np.array_equal(r.vertices, ref["verts"])


def test_device_filling_slabs_run_alone_in_the_mix():
    """Slabs of 2^21 vertices and more take the single-mesh pipeline inside a step
    while the small ones of the same step are batched; every result still equals
    the request processed alone, and the big request equals its one-shot
    extraction."""
    reqs = mixed_requests(6, seed=4)
    big = synth.gaussians((1100, 1000), 3, 12, seed=44)  # slab: 1100 x 1000 x 2 > 2^21 vertices
    reqs.append(serve.Request(len(reqs), 0.0, (1100, 1000), _lib.PRED_CRITICAL_POINT, 2, _lib.F64,
                              [[c[..., t] for c in big] for t in range(3)], simplices=1099 * 999 * 2 * 12,
                              meta={"comps": big}))
    srv = engine.Server(64 << 30)
    outs, wall, rej = serve.run(srv, reqs, keep_results=True)
    assert rej == [] and len(outs) == len(reqs)
    for o in outs:
        rq = reqs[o.rid]
        say(f"[serve] request {o.rid}: spatial={rq.spatial} T={len(rq.timesteps)} m={rq.m}")
        got = o.result.to_complex()
        same(got, alone(rq), f"request {o.rid}")
        if o.rid == len(reqs) - 1:
            assert len(got.vertices) > 100
            one = engine.extract(engine.RegularSimplicialMesh((1100, 1000, 3)), 0, 2, list(big)).to_complex()
            same(got, one, "big request vs one-shot")


def scalar_request(rid, kind, spatial, T, seed, dtype=np.float32, thr=(0.0, 0.0)):
    """A contour (one field) or fiber (two fields) stream request on a Gaussian sum
    minus its mean."""
    nf = 1 if kind == _lib.PRED_CONTOUR else 2
    fields = [synth.gaussians(spatial, T, 8, seed=seed + 17 * k, gradient=False, subtract_mean=True, dtype=dtype)[0]
              for k in range(nf)]
    m = 1 if kind == _lib.PRED_CONTOUR else 2
    dt = _lib.F32 if dtype == np.float32 else _lib.F64
    hc = int(np.prod([s - 1 for s in spatial])) * max(0, T - 1)
    return serve.Request(rid, 0.0, tuple(spatial), kind, m, dt, [[f[..., t] for f in fields] for t in range(T)],
                         thresholds=thr, simplices=hc, meta={"comps": fields})


def test_contour_and_fiber_requests_in_the_mix():
    """Contour and fiber streams (fiber on 3-D + time has star-fan cells, which are
    not batched and take the single-mesh pipeline) served beside critical-point
    streams: every result equals the request alone and its one-shot extraction."""
    reqs = mixed_requests(5, seed=6)
    n0 = len(reqs)
    reqs.append(scalar_request(n0, _lib.PRED_FIBER, (9, 8, 10), 4, 61, thr=(0.01, -0.02)))
    reqs.append(scalar_request(n0 + 1, _lib.PRED_FIBER, (24, 20), 5, 62))
    reqs.append(scalar_request(n0 + 2, _lib.PRED_CONTOUR, (12, 11, 10), 4, 63))
    reqs.append(scalar_request(n0 + 3, _lib.PRED_CONTOUR, (30, 26), 3, 64, dtype=np.float64, thr=(0.02, 0.0)))
    srv = engine.Server(64 << 30)
    outs, wall, rej = serve.run(srv, reqs, keep_results=True)
    assert rej == [] and len(outs) == len(reqs)
    for o in outs:
        rq = reqs[o.rid]
        say(f"[serve] request {o.rid}: kind={rq.kind} spatial={rq.spatial} T={len(rq.timesteps)} m={rq.m}")
        got = o.result.to_complex()
        same(got, alone(rq), f"request {o.rid}")
        if o.rid >= n0:
            comps = rq.meta["comps"]
            one = engine.extract(engine.RegularSimplicialMesh(list(comps[0].shape)), rq.kind, rq.m, comps,
                                 thresholds=rq.thresholds).to_complex()
            assert len(one.vertices) > 0
            same(got, one, f"request {o.rid} vs one-shot")


def test_several_device_filling_slabs_in_one_step():
    """Two or more single-mesh slabs per step: each slab's committed cells are
    re-indexed on its stream after the next slab's extraction, and the pool blocks of
    the slab they read are reused only once that has run.  Every result equals its
    one-shot extraction."""
    big = [synth.gaussians((1100, 1000), 4, 12, seed=70 + k) for k in range(2)]
    cube = synth.cp3d_config(n=(132, 128, 128), nt=3, ngauss=8, seed=72)
    reqs = [serve.Request(k, 0.0, (1100, 1000), _lib.PRED_CRITICAL_POINT, 2, _lib.F64,
                          [[c[..., t] for c in big[k]] for t in range(4)], simplices=1099 * 999 * 3 * 12,
                          meta={"comps": big[k]}) for k in range(2)]
    reqs.append(serve.Request(2, 0.0, (132, 128, 128), _lib.PRED_CRITICAL_POINT, 3, _lib.F64,
                              [[c[..., t] for c in cube] for t in range(3)], simplices=131 * 127 * 127 * 2 * 60,
                              meta={"comps": cube}))
    srv = engine.Server(64 << 30)
    outs, wall, rej = serve.run(srv, reqs, keep_results=True)
    assert rej == [] and len(outs) == 3
    for o in outs:
        rq = reqs[o.rid]
        comps = rq.meta["comps"]
        one = engine.extract(engine.RegularSimplicialMesh(list(comps[0].shape)), rq.kind, rq.m,
                             list(comps)).to_complex()
        got = o.result.to_complex()
        say(f"[serve] request {o.rid}: {len(got.vertices)} nodes")
        assert len(got.vertices) > 100
        same(got, one, f"request {o.rid} vs one-shot")


def test_noise_heavy_stream_stays_within_its_reservation():
    """A long, node-dense 3-D stream under a tight budget: the server admits it with
    the largest window of timesteps that fits, and served streams retire finished
    nodes to host memory past the retirement threshold, so once the server has
    measured the density (one short request of the same kind first), the device
    memory a stream holds stays within what admission reserved and below its whole
    payload; the result is unchanged."""
    T = 20
    comps = synth.cp3d_config(n=(34, 32, 30), nt=T, ngauss=6, seed=81, noise=0.05)
    steps = [[c[..., t] for c in comps] for t in range(T)]
    srv = engine.Server(48 << 20)
    # learn the density from a two-timestep request
    warm = srv.open((34, 32, 30), _lib.PRED_CRITICAL_POINT, 3, _lib.F64, 3, 2)
    for c in steps[:2]:
        srv.append(warm, c)
    srv.close_stream(warm)
    for _ in range(10):
        if srv.step():
            break
    srv.take(warm)
    assert srv.stats()["reserved_bytes"] == 0
    sid = srv.open((34, 32, 30), _lib.PRED_CRITICAL_POINT, 3, _lib.F64, 3, T)
    assert sid is not None
    reserved = srv.stats()["reserved_bytes"]
    for c in steps:
        srv.append(sid, c)
    srv.close_stream(sid)
    peak, done = 0, []
    for _ in range(2 * T):
        done = srv.step()
        peak = max(peak, srv.stats()["device_bytes"])
        if done:
            break
    assert done == [sid]
    got = srv.take(sid).to_complex()
    srv.close()
    payload = len(got.vertices) * (8 + 4 * 8 + 4 + 4 + 8)
    say(f"[serve] dense stream: {len(got.vertices)} nodes ({payload} B), device peak {peak} B, reserved {reserved} B")
    assert peak <= reserved  # the device holds no more than admission reserved...
    assert peak < payload  # ...and never the whole payload
    rq = serve.Request(0, 0.0, (34, 32, 30), _lib.PRED_CRITICAL_POINT, 3, _lib.F64, steps)
    same(got, alone(rq), "dense stream")
