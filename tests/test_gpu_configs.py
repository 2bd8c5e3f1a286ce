"""The batching and serving workloads at their real sizes, through the C ABI.

Batching: 1024 CriticalPointPredicate<2,double> jobs (sides log-uniform in [32, 1024],
2-16 timesteps, 4-64 moving Gaussians) in one batched pass under the autotuned
schedule; every job equals its own ftk2h_extract_regular bit for bit (the per-job
contract is execute(), engine.hpp:96), and a seeded sample of jobs, the smallest and
the largest included, equals the oracle.

Serving: Poisson arrivals of streaming requests (70% 2-D critical points at
128^2-2048^2, 30% 3-D at 64^3-256^3, 1-32 timesteps) under a 200 GB budget; every
served result equals the same request streamed alone (execute_stream,
engine.hpp:214-270), and a sample of 2-D requests equals the oracle's one-shot
extraction over the request's timesteps (the streaming contract,
test_examples.cpp:259-297)."""
import sys

import numpy as np
import pytest

import oracle
from ftk2_amd import _lib, engine, serve, synth

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b, what):
    assert np.array_equal(a.vertices, b.vertices), f"{what}: node simplices differ"
    assert np.array_equal(a.lam.view(np.uint64), b.lam.view(np.uint64)), f"{what}: lambda differs"
    assert np.array_equal(a.track_id, b.track_id), f"{what}: track ids differ"
    assert sorted(a.connectivity) == sorted(b.connectivity), f"{what}: connectivity dims differ"
    for d in a.connectivity:
        assert np.array_equal(a.connectivity[d], b.connectivity[d]), f"{what}: connectivity differs"


def host(t):
    """[x, y, (z), t] view of a device tensor -> Fortran-ordered numpy array."""
    return np.asfortranarray(t.cpu().numpy())


def lean_matches(cx, comps, m, what):
    L = oracle.Lean(comps, oracle.PRED_CP, m, nthreads=16)
    verts, lam, track = L.nodes()
    assert L.n_nodes == len(cx.vertices), f"{what}: node count"
    assert np.array_equal(cx.vertices, verts), f"{what}: nodes differ from the oracle"
    assert np.array_equal(cx.lam.view(np.uint64), lam.view(np.uint64)), f"{what}: lambda differs from the oracle"
    assert np.array_equal(cx.track_id, track), f"{what}: tracks differ from the oracle"
    if L.n_edges:
        assert np.array_equal(cx.connectivity[1], L.edges()), f"{what}: edges differ from the oracle"
    else:
        assert 1 not in cx.connectivity
    L.close()


def test_batching_1024_jobs_equal_plain_and_oracle():
    say("[config3] generating 1024 jobs")
    jobs = [(engine.RegularSimplicialMesh(list(dims)), [u, v]) for dims, u, v in synth.batch_jobs(1024, seed=0)]
    torch.cuda.synchronize()
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    cfg, ms = b.autotune(jobs, reps=1)
    say(f"[config3] autotuned schedule {cfg} at {ms:.1f} ms")
    res = b.run(jobs)
    nodes = 0
    for i, (mesh, comps) in enumerate(jobs):
        got = res[i].to_complex()
        want = engine.extract(mesh, _lib.PRED_CRITICAL_POINT, 2, comps).to_complex()
        same(got, want, f"job {i} {mesh.dims}")
        nodes += len(got.vertices)
        if (i + 1) % 256 == 0:
            say(f"[config3] {i + 1} jobs equal their plain extraction ({nodes} nodes so far)")
    assert nodes > 1000
    # oracle sample: smallest, largest and 6 seeded others
    size = [int(np.prod(m.dims)) for m, _ in jobs]
    rng = np.random.default_rng(11)
    pick = [int(np.argmin(size)), int(np.argmax(size))]
    pick += [int(i) for i in rng.choice([i for i in range(len(jobs)) if i not in pick], 6, replace=False)]
    for i in pick:
        mesh, comps = jobs[i]
        say(f"[config3] job {i} {mesh.dims} vs the oracle")
        lean_matches(res[i].to_complex(), [host(c) for c in comps], 2, f"job {i}")
    del res
    b.close()


def test_serving_mix_served_equals_alone_and_oracle():
    pools = synth.serving_pools(seed=0)
    torch.cuda.synchronize()
    rng = np.random.default_rng(5)
    shapes = synth.serving_mix(24, rng)
    reqs = []
    for i, (sd, n, T) in enumerate(shapes):
        perm = tuple(reversed(range(sd)))
        steps = [[p[t].permute(*perm) for p in pools[(sd, n)]] for t in range(T)]
        reqs.append(serve.Request(i, 0.0, (n,) * sd, _lib.PRED_CRITICAL_POINT, sd, _lib.F64, steps,
                                  simplices=(n - 1) ** sd * max(0, T - 1) * (12 if sd == 2 else 60),
                                  meta={"sd": sd, "n": n, "T": T}))
    assert {r.meta["sd"] for r in reqs} == {2, 3}
    arr = serve.poisson_arrivals(len(reqs), 40.0, np.random.default_rng(6))
    for rq, a in zip(reqs, arr):
        rq.arrival = float(a)
    srv = engine.Server(200 << 30)
    checked = []

    def on_result(rq, res):
        # compare as results arrive, so the device holds few results at once
        st = engine.Stream(rq.spatial, rq.kind, rq.m, rq.dtype, len(rq.timesteps[0]))
        for comps in rq.timesteps:
            st.append(comps)
        alone = st.finish().to_complex()
        st.close()
        same(res.to_complex(), alone, f"request {rq.rid} {rq.meta}")
        checked.append(rq.rid)

    outs, wall, rej = serve.run(srv, reqs, on_result=on_result)
    srv.close()
    assert rej == [] and len(outs) == len(reqs) and sorted(checked) == list(range(len(reqs)))
    lat = sorted(o.latency for o in outs)
    say(f"[config4] {len(outs)} requests served == alone in {wall:.2f} s, p50 {lat[len(lat) // 2] * 1e3:.1f} ms")
    # oracle: 2-D requests with at least two timesteps, smallest first, up to 4
    two_d = sorted((r for r in reqs if r.meta["sd"] == 2 and r.meta["T"] >= 2),
                   key=lambda r: (r.meta["n"] ** 2 * r.meta["T"], r.rid))[:4]
    assert two_d
    for rq in two_d:
        n, T = rq.meta["n"], rq.meta["T"]
        say(f"[config4] request {rq.rid} ({n}^2 x {T}) vs the oracle")
        comps = [np.asfortranarray(p[:T].permute(2, 1, 0).cpu().numpy()) for p in pools[(2, n)]]
        st = engine.Stream(rq.spatial, rq.kind, rq.m, rq.dtype, 2)
        for c in rq.timesteps:
            st.append(c)
        cx = st.finish().to_complex()
        st.close()
        lean_matches(cx, comps, 2, f"request {rq.rid}")
